// 16-column products of a structured node grid's 27-point block pattern:
// the FP64 Y = K X of the block Lanczos inner solves, the mass product of
// its outer loop, and the FP32 V-cycle passes (residual B - K X, damped
// block-Jacobi update) on 16 column-interleaved vectors, entry (3i + p, c)
// at (3i + p) * 16 + c.
//
// The pattern is the one a hex8 grid mesh gets when node id = ix + nx (iy +
// ny iz) and the elements are the grid's cells: row i holds its in-grid
// neighbours (|dx|, |dy|, |dz| <= 1) in (dz, dy, dx) order, so a row's start
// and the node of each of its blocks follow from the node's grid position;
// the kernel reads no index arrays.  (The mesh's pattern is checked against
// this structure's block count on entry.)
//
// One work-group owns an 8 x 4 tile of node columns and walks up a run of
// node planes, keeping the X rows of three planes of the tile and its
// one-node halo in an LDS ring (each plane fetched once, with coalesced
// 16-byte loads, one step ahead), so the neighbour gathers are LDS reads:
//   - thread (row r of the 32, lane j of 8) takes blocks j, j + 8, j + 16 and
//     j + 24 of its row and accumulates all 48 outputs of the row (3
//     components x 16 columns) from the neighbours' LDS rows; blocks past
//     the row's end read a zeroed LDS row (their K address is clamped to the
//     row's first block), so all four run without branches;
//   - lane j keeps column c in accumulator slot c ^ m_j (m_j from the bits
//     of j, below), so the 8 lanes of a row reduce-scatter the 48 sums by
//     three lane-move stages without per-lane selects, each lane ending with
//     2 columns x 3 components in a fixed order (deterministic); the swizzle
//     also spreads a read instruction's 8 lanes over the LDS banks;
//   - the K blocks of the next plane are loaded while the current one is
//     summed;
//   - the epilogue applies the mode: Y = K X, B - K X, or X + w Dinv (B - K X).
//
// Work-groups are dispatched round-robin over the 8 XCDs, each with its own
// L2; the launch gives each XCD a contiguous run of tiles, so the halo rows
// that neighbouring tiles share are fetched into that L2 once.
#include "common.hpp"

namespace fem {
namespace {

constexpr int kSt = 256;                                        // 32 rows x 8 lanes
constexpr int kStTX = 8, kStTY = 4;                             // node columns of a tile
constexpr int kStPX = kStTX + 2, kStPY = kStTY + 2;             // positions with halo
constexpr int kNb = 16;                                         // columns
constexpr int kNumXcd = 8;

// LDS ring geometry (values): a node's slot holds its 48 values; a halo row
// of kStPX slots plus RP - kStPX S padding; a plane of kStPY rows plus
// padding.  The paddings put the 8 lanes of a read instruction on distinct
// bank groups for interior rows (FP64 needs none with the column swizzle).
template <typename T> struct StGeom;
template <> struct StGeom<double> {
  using V = double2;
  static constexpr int S = 48, RP = kStPX * 48, PL = kStPY * RP;
};
template <> struct StGeom<float> {
  using V = float4;
  static constexpr int S = 48, RP = kStPX * 48 + 16, PL = kStPY * RP + 16;
};
// the ring's three planes, then one zeroed slot
template <typename T>
__host__ __device__ constexpr int st_lds_bytes() { return (3 * StGeom<T>::PL + StGeom<T>::S) * (int)sizeof(T); }

// Lane exchange by DPP within the 8 lanes of a row (64-bit values as two moves).
template <int CTRL>
__device__ __forceinline__ double st_dpp(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ float st_dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Barrier for LDS hand-offs only: the ring is the one thing the waves share,
// so outstanding global loads and stores need not drain first.
__device__ __forceinline__ void st_lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Makes v opaque at this point, so values derived from it are formed here
// rather than hoisted out of the plane loop (where they would hold registers
// across it).
__device__ __forceinline__ void st_opaque(int& v) { asm volatile("" : "+v"(v)); }
// Pins a block's sums before the next block's reads are issued (otherwise
// instruction selection may sink every FMA below all the LDS reads, holding
// their results in registers at once).
__device__ __forceinline__ void st_pin(double& v) { asm volatile("" : "+v"(v)); }
__device__ __forceinline__ void st_pin(float& v) { asm volatile("" : "+v"(v)); }

// Output stores: the 8 lanes of a row write whole 128-byte runs, so streaming
// (non-temporal) stores lose nothing and keep the outputs from evicting the
// X halo lines that neighbouring tiles still read (FP64 product 1.38 -> 1.35
// ms on the 500 x 60 x 75 blade).
template <typename T>
__device__ __forceinline__ void st_out(T* p, T v) { __builtin_nontemporal_store(v, p); }

// Blocks of the nodes before index n on an axis of N nodes (a node has
// 1 + (n > 0) + (n < N - 1) neighbours along the axis, itself included).
__device__ __forceinline__ int st_prefix(int n, int N) { return n == 0 ? 0 : (N > 1 ? 2 : 1) + 3 * (n - 1); }

// MODE 0: OUT = B - K X;  1: OUT = X + w Dinv (B - K X);  3: OUT = K X;
// 4: OUT = K X and part[work item][c] = sum over the item's nodes of
// X_c . (K X)_c (FP64: the p . q of conjugate gradients, in a fixed order).
// BW: values per block, 9 (3 x 3 blocks of K) or 1 (the mass matrix's one
// scalar per node pair on each DOF diagonal).
template <typename T, int MODE, int BW, int PF>
__global__ void __launch_bounds__(kSt, sizeof(T) == 8 ? 2 : 3)
stencil16_kernel(const T* __restrict__ vals, const T* __restrict__ X, const T* __restrict__ B,
                 const T* __restrict__ dinv, T w, T* __restrict__ OUT, int NX, int NY, int NZ, int seg, int ntx,
                 int nty, int nwork, int per_xcd, double* __restrict__ part) {
  using G = StGeom<T>;
  using V = typename G::V;
  constexpr int S = G::S, RP = G::RP, PL = G::PL;
  constexpr int CW = 16 / (int)sizeof(T);              // values per 16-byte chunk
  constexpr int NCU = kNb / CW;                        // chunks per component
  constexpr int CPN = 3 * NCU;                         // chunks per node
  constexpr int NCH = kStPX * kStPY * CPN;             // chunks per plane
  constexpr int LPT = (NCH + kSt - 1) / kSt;           // chunks per thread
  extern __shared__ __align__(16) unsigned char smem[];
  T* ring = reinterpret_cast<T*>(smem);
  T* zrow = ring + 3 * PL;

  // XCD x runs work items [x per_xcd, (x + 1) per_xcd) in order
  const int wid = (int)(blockIdx.x % kNumXcd) * per_xcd + (int)(blockIdx.x / kNumXcd);
  if (wid >= nwork) return;
  const int txi = wid % ntx, tyi = (wid / ntx) % nty, sg = wid / (ntx * nty);
  const int x0 = txi * kStTX, y0 = tyi * kStTY, z0 = sg * seg, z1 = min(z0 + seg, NZ);
  const int tid = threadIdx.x;
  const int r = tid >> 3, j = tid & 7;
  const int rx = x0 + (r & 7), ry = y0 + (r >> 3);
  const bool row_ok = rx < NX && ry < NY;
  const int plane = NX * NY;

  // column swizzle: slot s of a lane holds column s ^ mcol; partners of the
  // three exchange stages (j ^ 7, j ^ 2, j ^ 1) differ in mcol by 8, 4, 2
  const int b0 = j & 1, b1 = (j >> 1) & 1, b2 = j >> 2;
  const int mcol = sizeof(T) == 8 ? 8 * b2 + 4 * (b1 ^ b2) + 2 * (b0 ^ b2) : 8 * b2 + 4 * (b1 ^ b2);
  const int mu = mcol / CW;                            // the same on 16-byte chunks
  const int c0 = sizeof(T) == 8 ? mcol : mcol + 2 * b0;   // this lane's output columns c0, c0 + 1

  // the row's blocks j + 8q: plane step bz (0..2 from the lowest in-grid
  // plane) and in-plane ring offset; the row start follows the grid counts
  const int lox = rx > 0, loy = ry > 0;
  const int cx = 1 + lox + (rx < NX - 1), cy = 1 + loy + (ry < NY - 1), cxy = cx * cy;
  const int TX = 3 * NX - 2, TXY = TX * (3 * NY - 2);
  const int rowpart = st_prefix(ry, NY) * TX + cy * st_prefix(rx, NX);
  int rowp = rowpart, mus = mu;
  int pos[4], bzq[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int b = j + 8 * q;
    const int bz = b / cxy, rem = b - bz * cxy;
    const int by = rem / cx, bx = rem - by * cx;
    pos[q] = (ry - y0 + 1 + by - loy) * RP + (rx - x0 + 1 + bx - lox) * S;
    bzq[q] = bz;
  }
  auto kptr = [&](int z, int q) -> const T* {
    const int cz = 1 + (z > 0) + (z < NZ - 1);
    const int b = j + 8 * q;
    const bool ok = row_ok && b < cxy * cz;
    return vals + (ok ? (int64_t)BW * (st_prefix(z, NZ) * TXY + cz * rowp + b) : 0);
  };
  auto xptr = [&](int z, int q) -> const T* {
    const int lz = z > 0, cz = 1 + lz + (z < NZ - 1);
    const bool ok = row_ok && j + 8 * q < cxy * cz;
    int ri = (z - lz) % 3 + bzq[q];                    // ring slot of plane z - lz + bz
    ri = ri >= 3 ? ri - 3 : ri;
    return ok ? ring + ri * PL + pos[q] : zrow;
  };
  auto kload = [&](const T* kp, T (&k)[9]) {
    if (BW == 9) {
#pragma unroll
      for (int t = 0; t < 9; ++t) k[t] = kp[t];
    } else {
      k[0] = kp[0];
    }
  };
  auto pin_all = [](T (&acc)[3][kNb]) {
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int c = 0; c < kNb; ++c) st_pin(acc[p][c]);
  };
  auto accum = [&](const T* xs, const T (&k)[9], T (&acc)[3][kNb], bool first) {
#pragma unroll
    for (int u = 0; u < NCU; ++u) {
      const T* xp = xs + ((u ^ mus) * CW);
      const V v0 = *reinterpret_cast<const V*>(xp);
      const V v1 = *reinterpret_cast<const V*>(xp + kNb);
      const V v2 = *reinterpret_cast<const V*>(xp + 2 * kNb);
      const T* e0 = reinterpret_cast<const T*>(&v0);
      const T* e1 = reinterpret_cast<const T*>(&v1);
      const T* e2 = reinterpret_cast<const T*>(&v2);
#pragma unroll
      for (int h = 0; h < CW; ++h) {
        const int s = u * CW + h;
        if (BW == 9) {
#pragma unroll
          for (int p = 0; p < 3; ++p) {
            const T t2 = first ? k[3 * p + 2] * e2[h] : fma(k[3 * p + 2], e2[h], acc[p][s]);
            acc[p][s] = fma(k[3 * p], e0[h], fma(k[3 * p + 1], e1[h], t2));
          }
        } else {
          acc[0][s] = first ? k[0] * e0[h] : fma(k[0], e0[h], acc[0][s]);
          acc[1][s] = first ? k[0] * e1[h] : fma(k[0], e1[h], acc[1][s]);
          acc[2][s] = first ? k[0] * e2[h] : fma(k[0], e2[h], acc[2][s]);
        }
      }
      if (u == NCU / 2 - 1) pin_all(acc);     // half a block's reads in flight at a time
    }
  };

  // this thread's chunks of a plane: global offset within the plane (-1
  // outside the grid) and LDS offset within the ring plane
  int goff[LPT], loff[LPT];
#pragma unroll
  for (int i = 0; i < LPT; ++i) {
    const int e = tid + i * kSt;
    goff[i] = -1;
    loff[i] = 0;
    if (e < NCH) {
      const int p = e / CPN, k = e - p * CPN;
      const int px = p % kStPX, py = p / kStPX;
      const int gx = x0 - 1 + px, gy = y0 - 1 + py;
      loff[i] = py * RP + px * S + k * CW;
      if (gx >= 0 && gx < NX && gy >= 0 && gy < NY) goff[i] = (gx + NX * gy) * (3 * kNb) + k * CW;
    }
  }
  auto fetch = [&](int q, V (&v)[LPT]) {
    const bool in = q >= 0 && q < NZ;
    const T* Xq = X + (int64_t)plane * (in ? q : 0) * (3 * kNb);
#pragma unroll
    for (int i = 0; i < LPT; ++i) {
      v[i] = V{};
      if (in && goff[i] >= 0) v[i] = *reinterpret_cast<const V*>(Xq + goff[i]);
    }
  };
  auto stash = [&](int q, const V (&v)[LPT]) {
    T* dst = ring + (q % 3) * PL;
#pragma unroll
    for (int i = 0; i < LPT; ++i)
      if (tid + i * kSt < NCH) *reinterpret_cast<V*>(dst + loff[i]) = v[i];
  };

  if (tid < S / CW) *reinterpret_cast<V*>(zrow + tid * CW) = V{};
  V pre[LPT];
  double pq0 = 0.0, pq1 = 0.0;   // MODE 4: this lane's columns c0, c0 + 1
  if (z0 > 0) {
    fetch(z0 - 1, pre);
    stash(z0 - 1, pre);
  }
  fetch(z0, pre);
  stash(z0, pre);
  fetch(z0 + 1, pre);
  // K blocks in flight: PF buffers, block q in buffer q % PF, refilled with
  // block q + PF (the next plane's past 3) as soon as block q is summed
  T kq[PF][9];
#pragma unroll
  for (int q = 0; q < PF; ++q) kload(kptr(z0, q), kq[q]);

  for (int z = z0; z < z1; ++z) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      st_opaque(pos[q]);
      st_opaque(bzq[q]);
    }
#pragma unroll
    for (int i = 0; i < LPT; ++i) {
      st_opaque(goff[i]);
      st_opaque(loff[i]);
    }
    st_opaque(rowp);
    st_opaque(mus);
    // plane z+1 (fetched a step ago) into the slot of plane z-2, read by no one now
    stash(z + 1, pre);
    st_lds_barrier();
    fetch(z + 2, pre);

    const int inode = rx + NX * (ry + NY * z);
    const int64_t o = (int64_t)inode * (3 * kNb) + c0;
    T bv[3][2], dv[9];
    if (MODE < 3 && row_ok) {
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        bv[p][0] = B[o + p * kNb];
        bv[p][1] = B[o + p * kNb + 1];
      }
      if (MODE == 1) {
#pragma unroll
        for (int t = 0; t < 9; ++t) dv[t] = dinv[(int64_t)inode * 9 + t];
      }
    }

    // sum the four blocks (on the run's last plane the refills reload this
    // plane's blocks: no branch here, which would let the sums sink past all
    // the reads)
    T acc[3][kNb];
    const int zn = z + 1 < z1 ? z + 1 : z;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      accum(xptr(z, q), kq[q % PF], acc, q == 0);
      kload(q + PF < 4 ? kptr(z, q + PF) : kptr(zn, q + PF - 4), kq[q % PF]);
      pin_all(acc);
    }

    // reduce-scatter over the row's 8 lanes (no selects: see mcol)
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int s = 0; s < 8; ++s) acc[p][s] += st_dpp<0x141>(acc[p][s + 8]);   // row_half_mirror: j <-> 7 - j
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int s = 0; s < 4; ++s) acc[p][s] += st_dpp<0x4E>(acc[p][s + 4]);    // quad_perm [2,3,0,1]: j <-> j ^ 2
    if (sizeof(T) == 8) {
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int s = 0; s < 2; ++s) acc[p][s] += st_dpp<0xB1>(acc[p][s + 2]);  // quad_perm [1,0,3,2]: j <-> j ^ 1
    } else {
      // FP32 chunks hold 4 columns, so the last stage picks a half per lane
      const bool upper = b0 != 0;
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const T lo = acc[p][s], hi = acc[p][s + 2];
          acc[p][s] = (upper ? hi : lo) + st_dpp<0xB1>(upper ? lo : hi);
        }
    }
    // slots 0, 1 now hold columns c0, c0 + 1 of all three components
    if (row_ok) {
      if (MODE == 3 || MODE == 4) {
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          st_out(OUT + o + p * kNb, acc[p][0]);
          st_out(OUT + o + p * kNb + 1, acc[p][1]);
        }
        if (MODE == 4) {
          // this node's own X row, from the ring
          const T* xo = ring + (z % 3) * PL + (ry - y0 + 1) * RP + (rx - x0 + 1) * S + c0;
#pragma unroll
          for (int p = 0; p < 3; ++p) {
            pq0 = fma((double)xo[p * kNb], (double)acc[p][0], pq0);
            pq1 = fma((double)xo[p * kNb + 1], (double)acc[p][1], pq1);
          }
        }
      } else {
        T res[3][2];
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          res[p][0] = bv[p][0] - acc[p][0];
          res[p][1] = bv[p][1] - acc[p][1];
        }
        if (MODE == 0) {
#pragma unroll
          for (int p = 0; p < 3; ++p) {
            st_out(OUT + o + p * kNb, res[p][0]);
            st_out(OUT + o + p * kNb + 1, res[p][1]);
          }
        } else {
          // this node's own X row, from the ring
          const T* xo = ring + (z % 3) * PL + (ry - y0 + 1) * RP + (rx - x0 + 1) * S + c0;
#pragma unroll
          for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              const T d = w * (dv[3 * p] * res[0][h] + dv[3 * p + 1] * res[1][h] + dv[3 * p + 2] * res[2][h]);
              st_out(OUT + o + p * kNb + h, xo[p * kNb + h] + d);
            }
        }
      }
    }
    st_lds_barrier();   // every wave is done with plane z-1's slot before it takes plane z+2
  }
  if (MODE == 4) {
    // column c sits in slot c & 1 of the lane j whose c0 is c & ~1; sum the
    // 32 rows of the tile in row order (the ring is free after the last barrier)
    double* red = reinterpret_cast<double*>(smem);
    red[2 * tid] = pq0;
    red[2 * tid + 1] = pq1;
    __syncthreads();
    if (tid < kNb) {
      int jc = 0;
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const int a0 = jj & 1, a1 = (jj >> 1) & 1, a2 = jj >> 2;
        if (8 * a2 + 4 * (a1 ^ a2) + 2 * (a0 ^ a2) == (tid & ~1)) jc = jj;
      }
      double sum = 0.0;
      for (int rr = 0; rr < kSt / 8; ++rr) sum += red[2 * (rr * 8 + jc) + (tid & 1)];
      part[(int64_t)wid * kNb + tid] = sum;
    }
  }
}

int g_st_cus[64] = {0};

int st_cu_count() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (g_st_cus[dev] == 0) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    g_st_cus[dev] = n;
  }
  return g_st_cus[dev];
}

// Launch shape of a grid: 8 x 4 tiles times z segments, enough work-groups
// for two rounds over every CU's resident ones.
template <typename T>
struct StPlan {
  int ntx, nty, seg;
  int64_t nwork, per_xcd;
  StPlan(int64_t nx, int64_t ny, int64_t nz) {
    constexpr int lds = st_lds_bytes<T>();
    ntx = (int)ceil_div(nx, kStTX);
    nty = (int)ceil_div(ny, kStTY);
    const int64_t tiles = (int64_t)ntx * nty;
    const int64_t per_cu = std::min<int64_t>(sizeof(T) == 8 ? 2 : 3, (160 * 1024) / lds);
    const int64_t resident = (int64_t)st_cu_count() * per_cu;
    const int64_t nseg = std::max<int64_t>(1, std::min<int64_t>(nz, ceil_div(2 * resident, tiles)));
    seg = (int)ceil_div(nz, nseg);
    nwork = tiles * ceil_div(nz, seg);
    per_xcd = ceil_div(nwork, kNumXcd);
  }
};

template <typename T, int MODE, int BW = 9>
int launch_stencil16(int64_t nx, int64_t ny, int64_t nz, const T* vals, const T* X, const T* B, const T* dinv, T w,
                     T* OUT, hipStream_t st, double* part = nullptr, int64_t part_len = 0) {
  constexpr int lds = st_lds_bytes<T>();
  auto kern = stencil16_kernel<T, MODE, BW, (sizeof(T) == 8 && BW == 9) || MODE == 1 ? 2 : 4>;
  FEM_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  const StPlan<T> pl(nx, ny, nz);
  FEM_REQUIRE(pl.per_xcd * kNumXcd < ((int64_t)1 << 31), FEM_E_LIMIT, "grid too large for one launch");
  if (MODE == 4)
    FEM_REQUIRE(part && part_len >= pl.nwork * kNb, FEM_E_INVALID,
                "partials need %lld doubles (got %lld)", (long long)(pl.nwork * kNb), (long long)part_len);
  hipLaunchKernelGGL(kern, dim3((unsigned)(pl.per_xcd * kNumXcd)), dim3(kSt), lds, st, vals, X, B, dinv, w, OUT,
                     (int)nx, (int)ny, (int)nz, pl.seg, pl.ntx, pl.nty, (int)pl.nwork, (int)pl.per_xcd, part);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

int check_grid16(const fem_mesh* m, int64_t nx, int64_t ny, int64_t nz) {
  FEM_REQUIRE(m != nullptr, FEM_E_INVALID, "mesh is NULL");
  FEM_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && nx * ny * nz == m->num_node, FEM_E_INVALID,
              "node grid %lld x %lld x %lld does not match the mesh's %lld nodes", (long long)nx, (long long)ny,
              (long long)nz, (long long)m->num_node);
  FEM_REQUIRE(m->num_node * 3 * kNb < INT32_MAX, FEM_E_LIMIT, "48 x node count must fit int32");
  const int64_t blocks = (3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2);
  FEM_REQUIRE(m->nnzb == blocks, FEM_E_INVALID,
              "the mesh's %lld blocks are not the %lld of the grid's 27-point pattern", (long long)m->nnzb,
              (long long)blocks);
  FEM_REQUIRE(9 * blocks < INT32_MAX, FEM_E_LIMIT, "9 x block count must fit int32");
  return FEM_OK;
}

}  // namespace
}  // namespace fem

using namespace fem;

extern "C" int fem_bsr_spmm16_grid(const fem_mesh* mesh, int64_t nx, int64_t ny, int64_t nz, const double* vals,
                                   const double* X, double* Y, void* stream) {
  int rc = check_grid16(mesh, nx, ny, nz);
  if (rc) return rc;
  FEM_REQUIRE(vals && X && Y && X != Y, FEM_E_INVALID, "null pointer, or Y aliases X");
  FEM_REQUIRE(reinterpret_cast<uintptr_t>(X) % 16 == 0, FEM_E_INVALID, "X must be 16-byte aligned");
  return launch_stencil16<double, 3>(nx, ny, nz, vals, X, nullptr, nullptr, 0.0, Y, as_stream(stream));
}

extern "C" int fem_mass_spmm16_grid(const fem_mesh* mesh, int64_t nx, int64_t ny, int64_t nz,
                                    const double* m_blocks, const double* X, double* Y, void* stream) {
  int rc = check_grid16(mesh, nx, ny, nz);
  if (rc) return rc;
  FEM_REQUIRE(m_blocks && X && Y && X != Y, FEM_E_INVALID, "null pointer, or Y aliases X");
  FEM_REQUIRE(reinterpret_cast<uintptr_t>(X) % 16 == 0, FEM_E_INVALID, "X must be 16-byte aligned");
  return launch_stencil16<double, 3, 1>(nx, ny, nz, m_blocks, X, nullptr, nullptr, 0.0, Y, as_stream(stream));
}

extern "C" int fem_mg_smooth16_grid(const fem_mesh* mesh, int64_t nx, int64_t ny, int64_t nz, const float* vals,
                                    const float* dinv, float omega, const float* X, const float* B, float* OUT,
                                    int mode, void* stream) {
  int rc = check_grid16(mesh, nx, ny, nz);
  if (rc) return rc;
  FEM_REQUIRE(mode == 0 || mode == 1, FEM_E_INVALID, "mode must be 0 (residual) or 1 (update)");
  FEM_REQUIRE(vals && X && B && OUT && (mode == 0 || dinv), FEM_E_INVALID, "null pointer");
  FEM_REQUIRE(OUT != X, FEM_E_INVALID, "OUT must not alias X (X is gathered)");
  FEM_REQUIRE(reinterpret_cast<uintptr_t>(X) % 16 == 0, FEM_E_INVALID, "X must be 16-byte aligned");
  const hipStream_t st = as_stream(stream);
  if (mode == 0) return launch_stencil16<float, 0>(nx, ny, nz, vals, X, B, dinv, omega, OUT, st);
  return launch_stencil16<float, 1>(nx, ny, nz, vals, X, B, dinv, omega, OUT, st);
}

extern "C" int64_t fem_spmm16_grid_partials(int64_t nx, int64_t ny, int64_t nz) {
  if (nx < 1 || ny < 1 || nz < 1) return 0;
  return StPlan<double>(nx, ny, nz).nwork;
}

extern "C" int fem_bsr_spmm16_grid_pq(const fem_mesh* mesh, int64_t nx, int64_t ny, int64_t nz, const double* vals,
                                      const double* X, double* Y, double* part, int64_t part_len, void* stream) {
  int rc = check_grid16(mesh, nx, ny, nz);
  if (rc) return rc;
  FEM_REQUIRE(vals && X && Y && X != Y, FEM_E_INVALID, "null pointer, or Y aliases X");
  FEM_REQUIRE(reinterpret_cast<uintptr_t>(X) % 16 == 0, FEM_E_INVALID, "X must be 16-byte aligned");
  return launch_stencil16<double, 4>(nx, ny, nz, vals, X, nullptr, nullptr, 0.0, Y, as_stream(stream), part,
                                     part_len);
}
