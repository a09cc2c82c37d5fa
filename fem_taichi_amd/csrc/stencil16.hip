// 16-column products of a structured node grid's 27-point block pattern:
// the FP64 Y = K X of the block Lanczos inner solves and the FP32 V-cycle
// passes (residual B - K X, damped block-Jacobi update) on 16
// column-interleaved vectors, entry (3i + p, c) at (3i + p) * 16 + c.
//
// The pattern is the one of a structured grid, node id = ix + nx (iy + ny iz):
// every column of a row lies within one index step along each axis.  With a
// pattern-agnostic kernel every lane gathers the 27 neighbours' X rows through
// the vector memory pipe (27 x 384 bytes per row at FP64, over five times the
// row's K blocks); that pipe, not HBM, set the time.  Here one work-group
// owns an 8 x 4 tile of node columns and walks up a run of node planes,
// keeping the X rows of three planes of the tile and its one-node halo in an
// LDS ring (each plane fetched once, with coalesced 16-byte loads, one step
// ahead), so the gathers become LDS reads:
//   - thread (row r of the 32, lane j of 8) takes blocks j, j + 8, j + 16 and
//     j + 24 of its row (coalesced 16-byte loads of the K values, the column
//     index from ReNode) and accumulates all 48 outputs of the row (3
//     components x 16 columns) from the neighbours' LDS rows;
//   - the 8 lanes of a row then reduce-scatter the 48 sums by lane moves
//     (halves, quarters, eighths of the columns), each lane ending with 2
//     columns x 3 components in a fixed order (deterministic);
//   - the epilogue applies the mode: Y = K X, B - K X, or X + w Dinv (B - K X).
#include "common.hpp"

namespace fem {
namespace {

constexpr int kSt = 256;                                        // 32 rows x 8 lanes
constexpr int kStTX = 8, kStTY = 4;                             // node columns of a tile
constexpr int kStPX = kStTX + 2, kStPY = kStTY + 2, kStNP = kStPX * kStPY;   // positions with halo
constexpr int kNb = 16;                                         // columns

// LDS row of one node: 48 values padded to an odd number of 16-byte chunks,
// so consecutive positions start on different bank groups.
template <typename T>
__host__ __device__ constexpr int st_slot() { return sizeof(T) == 8 ? 50 : 52; }
template <typename T>
__host__ __device__ constexpr int st_lds_bytes() { return 3 * kStNP * st_slot<T>() * (int)sizeof(T); }

template <typename T> struct Vec16;
template <> struct Vec16<double> { using type = double2; };
template <> struct Vec16<float> { using type = float4; };

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

// One reduce-scatter stage over the lanes paired by CTRL: of the W column
// groups a lane holds, the lanes with `upper` keep the upper half, the
// others the lower half, each adding its partner's copy of what it keeps.
template <typename T, int CTRL, int W>
__device__ __forceinline__ void st_stage(T (&a)[3][kNb], bool upper) {
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int c = 0; c < W / 2; ++c) {
      const T lo = a[p][c], hi = a[p][c + W / 2];
      const T send = upper ? lo : hi;
      const T keep = upper ? hi : lo;
      a[p][c] = keep + st_dpp<CTRL>(send);
    }
}

// MODE 0: OUT = B - K X;  1: OUT = X + w Dinv (B - K X);  3: OUT = K X.
// BW: values per block, 9 (3 x 3 blocks of K) or 1 (the mass matrix's one
// scalar per node pair on each DOF diagonal).
template <typename T, int MODE, int BW = 9>
__global__ void __launch_bounds__(kSt, sizeof(T) == 8 ? 2 : 3)
stencil16_kernel(const int32_t* __restrict__ sm_node, const int32_t* __restrict__ co_node,
                 const int32_t* __restrict__ re_node, const T* __restrict__ vals, const T* __restrict__ X,
                 const T* __restrict__ B, const T* __restrict__ dinv, T w, T* __restrict__ OUT, int NX, int NY,
                 int NZ, int seg, int ntx, int nty) {
  using V = typename Vec16<T>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  T* ring = reinterpret_cast<T*>(smem);                // [3][kStNP][slot]
  constexpr int S = st_slot<T>();
  constexpr int CW = 16 / (int)sizeof(T);              // values per 16-byte chunk
  constexpr int CPN = 3 * kNb / CW;                    // chunks per node
  constexpr int NCH = kStNP * CPN;                     // chunks per plane
  constexpr int LPT = (NCH + kSt - 1) / kSt;           // chunks per thread

  const int blk = blockIdx.x;
  const int txi = blk % ntx, tyi = (blk / ntx) % nty, sg = blk / (ntx * nty);
  const int x0 = txi * kStTX, y0 = tyi * kStTY, z0 = sg * seg, z1 = min(z0 + seg, NZ);
  const int tid = threadIdx.x;
  const int r = tid >> 3, j = tid & 7;
  const int rx = x0 + (r & 7), ry = y0 + (r >> 3);
  const bool row_ok = rx < NX && ry < NY;
  const int64_t plane = (int64_t)NX * NY;
  const int plane32 = NX * NY;

  auto fetch = [&](int q, V (&v)[LPT]) {
#pragma unroll
    for (int i = 0; i < LPT; ++i) {
      v[i] = V{};
      const int e = tid + i * kSt;
      if (e < NCH && q >= 0 && q < NZ) {
        const int pos = e / CPN, k = e - pos * CPN;
        const int gx = x0 - 1 + pos % kStPX, gy = y0 - 1 + pos / kStPX;
        if (gx >= 0 && gx < NX && gy >= 0 && gy < NY)
          v[i] = *reinterpret_cast<const V*>(X + (gx + (int64_t)NX * gy + plane * q) * (3 * kNb) + k * CW);
      }
    }
  };
  auto stash = [&](int q, const V (&v)[LPT]) {
#pragma unroll
    for (int i = 0; i < LPT; ++i) {
      const int e = tid + i * kSt;
      if (e < NCH) {
        const int pos = e / CPN, k = e - pos * CPN;
        *reinterpret_cast<V*>(ring + ((q % 3) * kStNP + pos) * S + k * CW) = v[i];
      }
    }
  };

  V pre[LPT];
  if (z0 > 0) {
    fetch(z0 - 1, pre);
    stash(z0 - 1, pre);
  }
  fetch(z0, pre);
  stash(z0, pre);
  fetch(z0 + 1, pre);
  for (int z = z0; z < z1; ++z) {
    // plane z+1 (fetched a step ago) into the slot of plane z-2, read by no one now
    stash(z + 1, pre);
    __syncthreads();
    fetch(z + 2, pre);

    T acc[3][kNb];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int c = 0; c < kNb; ++c) acc[p][c] = T(0);
    const int64_t i = rx + (int64_t)NX * ry + plane * z;
    if (row_ok) {
      const int64_t sm = sm_node[i];
      const int co = co_node[i];
      // neighbour slot base of each block, then the blocks
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int b = j + 8 * q;
        if (b < co) {
          const int col = re_node[sm + b];
          const int cz = col / plane32;
          const int rem = col - cz * plane32;
          const int cy = rem / NX, cx = rem - cy * NX;
          const T* xs = ring + ((cz % 3) * kStNP + (cy - y0 + 1) * kStPX + (cx - x0 + 1)) * S;
          const T* kv = vals + BW * (sm + b);
          T kb[9];
#pragma unroll
          for (int t = 0; t < 9; ++t) kb[t] = BW == 9 ? kv[t] : (t % 4 == 0 ? kv[0] : T(0));
#pragma unroll
          for (int c = 0; c < kNb; c += CW) {
            const V v0 = *reinterpret_cast<const V*>(xs + c);
            const V v1 = *reinterpret_cast<const V*>(xs + kNb + c);
            const V v2 = *reinterpret_cast<const V*>(xs + 2 * kNb + c);
            const T* e0 = reinterpret_cast<const T*>(&v0);
            const T* e1 = reinterpret_cast<const T*>(&v1);
            const T* e2 = reinterpret_cast<const T*>(&v2);
#pragma unroll
            for (int u = 0; u < CW; ++u) {
              if (BW == 9) {
#pragma unroll
                for (int p = 0; p < 3; ++p)
                  acc[p][c + u] =
                      fma(kb[3 * p], e0[u], fma(kb[3 * p + 1], e1[u], fma(kb[3 * p + 2], e2[u], acc[p][c + u])));
              } else {
                acc[0][c + u] = fma(kb[0], e0[u], acc[0][c + u]);
                acc[1][c + u] = fma(kb[0], e1[u], acc[1][c + u]);
                acc[2][c + u] = fma(kb[0], e2[u], acc[2][c + u]);
              }
            }
          }
        }
      }
    }
    // reduce-scatter over the row's 8 lanes: lane j ends with columns 2*j', 2*j'+1
    // (j' the lane's position in the order below) of all three components
    st_stage<T, 0x141, 16>(acc, j >= 4);             // row_half_mirror: j <-> 7 - j
    st_stage<T, 0x4E, 8>(acc, (j & 2) != 0);         // quad_perm [2,3,0,1]: j <-> j ^ 2
    st_stage<T, 0xB1, 4>(acc, (j & 1) != 0);         // quad_perm [1,0,3,2]: j <-> j ^ 1
    // columns held: 8 (j >= 4) + 4 ((j & 2) != 0) + 2 ((j & 1) != 0), plus 0, 1
    const int c0 = 8 * (j >= 4) + 4 * ((j & 2) != 0) + 2 * (j & 1);
    if (row_ok) {
      const int64_t o = 3 * i * kNb + c0;
      if (MODE == 3) {
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          OUT[o + p * kNb] = acc[p][0];
          OUT[o + p * kNb + 1] = acc[p][1];
        }
      } else {
        T res[3][2];
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          res[p][0] = B[o + p * kNb] - acc[p][0];
          res[p][1] = B[o + p * kNb + 1] - acc[p][1];
        }
        if (MODE == 0) {
#pragma unroll
          for (int p = 0; p < 3; ++p) {
            OUT[o + p * kNb] = res[p][0];
            OUT[o + p * kNb + 1] = res[p][1];
          }
        } else {
          const T* D = dinv + 9 * i;
#pragma unroll
          for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              const T d = w * (D[3 * p] * res[0][h] + D[3 * p + 1] * res[1][h] + D[3 * p + 2] * res[2][h]);
              OUT[o + p * kNb + h] = X[o + p * kNb + h] + d;
            }
        }
      }
    }
    __syncthreads();   // every wave is done with plane z-1's slot before it takes plane z+2
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

template <typename T, int MODE, int BW = 9>
int launch_stencil16(const fem_mesh& m, int64_t nx, int64_t ny, int64_t nz, const T* vals, const T* X, const T* B,
                     const T* dinv, T w, T* OUT, hipStream_t st) {
  constexpr int lds = st_lds_bytes<T>();
  auto kern = stencil16_kernel<T, MODE, BW>;
  FEM_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  const int ntx = (int)ceil_div(nx, kStTX), nty = (int)ceil_div(ny, kStTY);
  const int64_t tiles = (int64_t)ntx * nty;
  // z segments: enough work-groups for two rounds over every CU's resident ones
  const int64_t resident = (int64_t)st_cu_count() * ((160 * 1024) / lds);
  const int64_t nseg = std::max<int64_t>(1, std::min<int64_t>(nz, ceil_div(2 * resident, tiles)));
  const int seg = (int)ceil_div(nz, nseg);
  const int64_t blocks = tiles * ceil_div(nz, seg);
  FEM_REQUIRE(blocks < ((int64_t)1 << 31), FEM_E_LIMIT, "grid too large for one launch");
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(kSt), lds, st, m.sm_node, m.co_node, m.re_node, vals, X, B,
                     dinv, w, OUT, (int)nx, (int)ny, (int)nz, seg, ntx, nty);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

int check_grid16(const fem_mesh* m, int64_t nx, int64_t ny, int64_t nz) {
  FEM_REQUIRE(m != nullptr, FEM_E_INVALID, "mesh is NULL");
  FEM_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && nx * ny * nz == m->num_node, FEM_E_INVALID,
              "node grid %lld x %lld x %lld does not match the mesh's %lld nodes", (long long)nx, (long long)ny,
              (long long)nz, (long long)m->num_node);
  FEM_REQUIRE(m->num_node < INT32_MAX, FEM_E_LIMIT, "node ids must fit int32");
  FEM_REQUIRE(m->sm_node && m->co_node && m->re_node, FEM_E_INVALID, "pattern missing");
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
  return launch_stencil16<double, 3>(*mesh, nx, ny, nz, vals, X, nullptr, nullptr, 0.0, Y, as_stream(stream));
}

extern "C" int fem_mass_spmm16_grid(const fem_mesh* mesh, int64_t nx, int64_t ny, int64_t nz,
                                    const double* m_blocks, const double* X, double* Y, void* stream) {
  int rc = check_grid16(mesh, nx, ny, nz);
  if (rc) return rc;
  FEM_REQUIRE(m_blocks && X && Y && X != Y, FEM_E_INVALID, "null pointer, or Y aliases X");
  FEM_REQUIRE(reinterpret_cast<uintptr_t>(X) % 16 == 0, FEM_E_INVALID, "X must be 16-byte aligned");
  return launch_stencil16<double, 3, 1>(*mesh, nx, ny, nz, m_blocks, X, nullptr, nullptr, 0.0, Y, as_stream(stream));
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
  if (mode == 0) return launch_stencil16<float, 0>(*mesh, nx, ny, nz, vals, X, B, dinv, omega, OUT, st);
  return launch_stencil16<float, 1>(*mesh, nx, ny, nz, vals, X, B, dinv, omega, OUT, st);
}
