// BSR(3x3) SpMV / SpMM, block-Jacobi, and the fused PCG iteration kernels.
//
// SpMV: a group of 32 lanes owns one block row; lane b loads block b's nine
// values (the row's blocks are contiguous, so the group reads one contiguous
// 72*CoNode-byte run), its column index and the three x entries of that
// column, and the three row sums are reduced across the 32 lanes with
// shuffles.  Grids are fixed (kGrid blocks, grid-stride over rows) so that
// dot products are reduced without atomics: each of exactly kGrid blocks
// writes one partial (zero when it had no rows),
// and the kernel that needs the scalar re-sums the kGrid partials itself in
// index order (every block gets the identical, deterministic value).
#include "common.hpp"

namespace fem {
namespace {

constexpr int kRowLanes = 32;
constexpr int kThreads = 256;
constexpr int kRowsPerBlock = kThreads / kRowLanes;   // 8
constexpr int kGrid = 1024;                            // blocks of every reducing kernel

// Sum of a double over the block (fixed tree); every thread gets the result.
__device__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  const int nw = blockDim.x >> 6;
  for (int q = 0; q < nw; ++q) s += red[q];
  return s;
}

// Deterministic sum of kGrid partials (same order in every block).
__device__ double sum_partials(const double* __restrict__ part, double* red) {
  double s = 0.0;
  for (int b = threadIdx.x; b < kGrid; b += blockDim.x) s += part[b];
  return block_sum(s, red);
}

// dst[b] = sum of src[b*chunk .. (b+1)*chunk) for the kGrid blocks b, chunk =
// ceil(n / kGrid) (fixed order: deterministic).
__global__ void __launch_bounds__(kThreads)
fold_partials_kernel(const double* __restrict__ src, int64_t n, double* __restrict__ dst) {
  __shared__ double red[kThreads / 64];
  const int64_t chunk = (n + kGrid - 1) / kGrid;
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < n ? lo + chunk : n;
  double s = 0.0;
  for (int64_t j = lo + threadIdx.x; j < hi; j += blockDim.x) s += src[j];
  s = block_sum(s, red);
  if (threadIdx.x == 0) dst[blockIdx.x] = s;
}

template <typename VT, typename XT>
__device__ __forceinline__ void row_product(const fem_mesh& m, const VT* __restrict__ vals,
                                            const XT* __restrict__ x, int64_t i, int lane,
                                            XT& a0, XT& a1, XT& a2) {
  a0 = a1 = a2 = XT(0);
  const int64_t sm = m.sm_node[i];
  const int co = m.co_node[i];
#pragma unroll 2
  for (int b = lane; b < co; b += kRowLanes) {
    const VT* __restrict__ v = vals + 9 * (sm + b);
    const int64_t c = 3 * (int64_t)m.re_node[sm + b];
    const XT x0 = x[c], x1 = x[c + 1], x2 = x[c + 2];
    a0 += (XT)v[0] * x0 + (XT)v[1] * x1 + (XT)v[2] * x2;
    a1 += (XT)v[3] * x0 + (XT)v[4] * x1 + (XT)v[5] * x2;
    a2 += (XT)v[6] * x0 + (XT)v[7] * x1 + (XT)v[8] * x2;
  }
#pragma unroll
  for (int o = kRowLanes / 2; o > 0; o >>= 1) {
    a0 += __shfl_xor(a0, o, 64);
    a1 += __shfl_xor(a1, o, 64);
    a2 += __shfl_xor(a2, o, 64);
  }
}

// y = A x (one row per 32-lane team).  With DOT, partial x.y per block into
// part[blockIdx.x]; block 0 also moves state[1] (rz of the previous
// iteration) into state[0] for the update kernel that follows.
template <typename VT, typename XT, bool DOT, int NT = kThreads>
__global__ void __launch_bounds__(NT)
spmv_kernel(fem_mesh m, const VT* __restrict__ vals, const XT* __restrict__ x, XT* __restrict__ y,
            double* __restrict__ part, double* state) {
  __shared__ double red[NT / 64];
  constexpr int kRows = NT / kRowLanes;
  const int lane = threadIdx.x & (kRowLanes - 1);
  double d = 0.0;
  // one row per lane team: the grid covers every row (the launch sites use ceil(N / kRows) blocks)
  const int64_t i = (int64_t)blockIdx.x * kRows + threadIdx.x / kRowLanes;
  if (i < m.num_node) {
    XT a0, a1, a2;
    row_product(m, vals, x, i, lane, a0, a1, a2);
    if (lane == 0) {
      y[3 * i] = a0;
      y[3 * i + 1] = a1;
      y[3 * i + 2] = a2;
      if (DOT) d += (double)a0 * (double)x[3 * i] + (double)a1 * (double)x[3 * i + 1] +
                    (double)a2 * (double)x[3 * i + 2];
    }
  }
  if (DOT) {
    d = block_sum(d, red);
    if (threadIdx.x == 0) {
      part[blockIdx.x] = d;
      if (blockIdx.x == 0) state[0] = state[1];
    }
  }
}

// SpMM: Y = A X, X/Y (3N, nrhs) row-major.  Thread per (block row, column).
__global__ void spmm_kernel(fem_mesh m, const double* __restrict__ vals, const double* __restrict__ X,
                            double* __restrict__ Y, int nrhs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t / nrhs;
  const int c = (int)(t - i * nrhs);
  if (i >= m.num_node) return;
  const int64_t sm = m.sm_node[i];
  const int co = m.co_node[i];
  double a0 = 0, a1 = 0, a2 = 0;
#pragma unroll 4
  for (int b = 0; b < co; ++b) {
    const double* v = vals + 9 * (sm + b);
    const int64_t col = m.re_node[sm + b];
    const double x0 = X[(3 * col) * nrhs + c], x1 = X[(3 * col + 1) * nrhs + c],
                 x2 = X[(3 * col + 2) * nrhs + c];
    a0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
    a1 += v[3] * x0 + v[4] * x1 + v[5] * x2;
    a2 += v[6] * x0 + v[7] * x1 + v[8] * x2;
  }
  Y[(3 * i) * nrhs + c] = a0;
  Y[(3 * i + 1) * nrhs + c] = a1;
  Y[(3 * i + 2) * nrhs + c] = a2;
}

// SpMM with two columns per thread (nrhs even, >= 8): each K block is loaded
// once for the pair and the X rows are read as 16-byte pairs.  Measured on
// the 1M-node blade: 16 columns 2.20 -> 1.36 ms, 8 columns 1.25 -> 1.10 ms;
// at 4 columns the thread-per-column form is faster.  Same sums in the same
// order as spmm_kernel.
__global__ void spmm2_kernel(fem_mesh m, const double* __restrict__ vals, const double* __restrict__ X,
                             double* __restrict__ Y, int nrhs) {
  const int npair = nrhs >> 1;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t / npair;
  const int c = 2 * (int)(t - i * npair);
  if (i >= m.num_node) return;
  const int64_t sm = m.sm_node[i];
  const int co = m.co_node[i];
  double a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0;
#pragma unroll 4   // more gathers in flight (2.38 -> 2.21 ms at 8 columns on 2.3M nodes)
  for (int b = 0; b < co; ++b) {
    const double* v = vals + 9 * (sm + b);
    const int64_t col = m.re_node[sm + b];
    const double2 x0 = *reinterpret_cast<const double2*>(X + (3 * col) * nrhs + c);
    const double2 x1 = *reinterpret_cast<const double2*>(X + (3 * col + 1) * nrhs + c);
    const double2 x2 = *reinterpret_cast<const double2*>(X + (3 * col + 2) * nrhs + c);
    a0 += v[0] * x0.x + v[1] * x1.x + v[2] * x2.x;
    a1 += v[3] * x0.x + v[4] * x1.x + v[5] * x2.x;
    a2 += v[6] * x0.x + v[7] * x1.x + v[8] * x2.x;
    b0 += v[0] * x0.y + v[1] * x1.y + v[2] * x2.y;
    b1 += v[3] * x0.y + v[4] * x1.y + v[5] * x2.y;
    b2 += v[6] * x0.y + v[7] * x1.y + v[8] * x2.y;
  }
  *reinterpret_cast<double2*>(Y + (3 * i) * nrhs + c) = make_double2(a0, b0);
  *reinterpret_cast<double2*>(Y + (3 * i + 1) * nrhs + c) = make_double2(a1, b1);
  *reinterpret_cast<double2*>(Y + (3 * i + 2) * nrhs + c) = make_double2(a2, b2);
}

// SpMM for 16 columns: the 8 lanes of a row (two columns each) load the row's
// K blocks eight at a time with coalesced 8-byte loads into the wave's LDS
// slice (9 loads per lane per 8 blocks instead of 9 per block), then read
// each block from LDS.  The wave's lanes only exchange data with lanes of
// their own row, so a wave-level fence orders the LDS writes and reads.
// Same sums in the same order as spmm2_kernel (bit-identical); on the
// 500x60x75 blade 2.56 ms instead of 2.80 ms (at 8 columns the plain kernel
// is faster).
constexpr int kS16Team = 8, kS16Rows = 256 / kS16Team / 4, kS16Stride = 9 * kS16Team + 1;
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__global__ void __launch_bounds__(256)
spmm16_kernel(fem_mesh m, const double* __restrict__ vals, const double* __restrict__ X, double* __restrict__ Y) {
  constexpr int nrhs = 16;
  __shared__ double sk[4][kS16Rows][kS16Stride];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int r = l / kS16Team, j = l % kS16Team;
  const int64_t i = ((int64_t)blockIdx.x * 256 + threadIdx.x) / kS16Team;
  if (i >= m.num_node) return;       // whole row teams leave together
  const int64_t sm = m.sm_node[i];
  const int co = m.co_node[i];
  const int c = 2 * j;
  double a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0;
  double* s = &sk[w][r][0];
  for (int c0 = 0; c0 < co; c0 += kS16Team) {
    const int nblk = min(kS16Team, co - c0);
    const double* src = vals + 9 * (sm + c0);
    double kv[9];                    // the nine loads in flight together, then the LDS stores
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int d = j + kS16Team * k;
      kv[k] = d < 9 * nblk ? src[d] : double(0);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int d = j + kS16Team * k;
      if (d < 9 * nblk) s[d] = kv[k];
    }
    const int colj = j < nblk ? m.re_node[sm + c0 + j] : 0;
    wave_lds_sync();
#pragma unroll
    for (int b = 0; b < kS16Team; ++b) {
      if (b < nblk) {
        const double* v = s + 9 * b;
        const int64_t col = __shfl(colj, b, kS16Team);
        const double2 x0 = *reinterpret_cast<const double2*>(X + (3 * col) * nrhs + c);
        const double2 x1 = *reinterpret_cast<const double2*>(X + (3 * col + 1) * nrhs + c);
        const double2 x2 = *reinterpret_cast<const double2*>(X + (3 * col + 2) * nrhs + c);
        a0 += v[0] * x0.x + v[1] * x1.x + v[2] * x2.x;
        a1 += v[3] * x0.x + v[4] * x1.x + v[5] * x2.x;
        a2 += v[6] * x0.x + v[7] * x1.x + v[8] * x2.x;
        b0 += v[0] * x0.y + v[1] * x1.y + v[2] * x2.y;
        b1 += v[3] * x0.y + v[4] * x1.y + v[5] * x2.y;
        b2 += v[6] * x0.y + v[7] * x1.y + v[8] * x2.y;
      }
    }
    wave_lds_sync();    // this chunk's readers are done before the next chunk overwrites the slice
  }
  *reinterpret_cast<double2*>(Y + (3 * i) * nrhs + c) = make_double2(a0, b0);
  *reinterpret_cast<double2*>(Y + (3 * i + 1) * nrhs + c) = make_double2(a1, b1);
  *reinterpret_cast<double2*>(Y + (3 * i + 2) * nrhs + c) = make_double2(a2, b2);
}

// Mass product with one scalar per node-pair block on each DOF diagonal.
__global__ void mass_spmm_kernel(fem_mesh m, const double* __restrict__ mb, const double* __restrict__ X,
                                 double* __restrict__ Y, int nrhs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t / nrhs;
  const int c = (int)(t - i * nrhs);
  if (i >= m.num_node) return;
  const int64_t sm = m.sm_node[i];
  const int co = m.co_node[i];
  double a0 = 0, a1 = 0, a2 = 0;
#pragma unroll 4
  for (int b = 0; b < co; ++b) {
    const double w = mb[sm + b];
    const int64_t col = m.re_node[sm + b];
    a0 += w * X[(3 * col) * nrhs + c];
    a1 += w * X[(3 * col + 1) * nrhs + c];
    a2 += w * X[(3 * col + 2) * nrhs + c];
  }
  Y[(3 * i) * nrhs + c] = a0;
  Y[(3 * i + 1) * nrhs + c] = a1;
  Y[(3 * i + 2) * nrhs + c] = a2;
}

// Inverse of each node's diagonal 3x3 block (FP64 arithmetic).
template <typename T>
__global__ void bj_setup_kernel(fem_mesh m, const T* __restrict__ vals, T* __restrict__ dinv,
                                int32_t* flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.num_node) return;
  const int64_t sm = m.sm_node[i];
  const int co = m.co_node[i];
  int k = -1;
  for (int q = 0; q < co; ++q)
    if (m.re_node[sm + q] == i) { k = q; break; }
  double B[9], Bi[9];
  bool ok = k >= 0;
  if (ok) {
    for (int q = 0; q < 9; ++q) B[q] = (double)vals[9 * (sm + k) + q];
    const double det = inv3(B, Bi);
    ok = (det != 0.0) && isfinite(det);
    for (int q = 0; q < 9; ++q) ok = ok && isfinite(Bi[q]);
  }
  if (!ok) {
    for (int q = 0; q < 9; ++q) Bi[q] = (q % 4 == 0) ? 1.0 : 0.0;
    if (flag) atomicMax(flag, 1);
  }
  for (int q = 0; q < 9; ++q) dinv[9 * i + q] = (T)Bi[q];
}

// r = b - q; z = Dinv r; p = z; partials of r.z, r.r, b.b (q may alias z:
// each thread reads its own entries before writing them).
template <typename VT, typename XT>
__global__ void __launch_bounds__(kThreads)
pcg_start_kernel(int64_t N, const XT* __restrict__ b, const XT* q, const VT* __restrict__ dinv,
                 XT* r, XT* z, XT* p, double* __restrict__ part) {
  __shared__ double red[kThreads / 64];
  double rz = 0, rr = 0, bb = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    XT ri[3], zi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) ri[a] = b[3 * i + a] - q[3 * i + a];
    const VT* D = dinv + 9 * i;
#pragma unroll
    for (int a = 0; a < 3; ++a)
      zi[a] = (XT)D[3 * a] * ri[0] + (XT)D[3 * a + 1] * ri[1] + (XT)D[3 * a + 2] * ri[2];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      r[3 * i + a] = ri[a];
      z[3 * i + a] = zi[a];
      p[3 * i + a] = zi[a];
      rz += (double)ri[a] * (double)zi[a];
      rr += (double)ri[a] * (double)ri[a];
      const double bv = (double)b[3 * i + a];
      bb += bv * bv;
    }
  }
  rz = block_sum(rz, red);
  rr = block_sum(rr, red);
  bb = block_sum(bb, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = rz;
    part[kGrid + blockIdx.x] = rr;
    part[2 * kGrid + blockIdx.x] = bb;
  }
}

// One block: state[1] = rz (as "next"), state[4] = rr, state[5] = bb.
__global__ void pcg_start_finalize(const double* __restrict__ part, double* state, int with_bb) {
  __shared__ double red[kThreads / 64];
  const double rz = sum_partials(part, red);
  const double rr = sum_partials(part + kGrid, red);
  const double bb = sum_partials(part + 2 * kGrid, red);
  if (threadIdx.x == 0) {
    state[1] = rz;
    state[4] = rr;
    if (with_bb) state[5] = bb;
  }
}

// alpha = state[0] / sum(part_pq); x += alpha p; r -= alpha q; z = Dinv r;
// partials of r.z and r.r.
template <typename VT, typename XT>
__global__ void __launch_bounds__(kThreads)
pcg_update_kernel(int64_t N, XT* __restrict__ x, XT* __restrict__ r, XT* __restrict__ z,
                  const XT* __restrict__ p, const XT* __restrict__ q, const VT* __restrict__ dinv,
                  const double* __restrict__ part_pq, double* __restrict__ part_out, double* state) {
  __shared__ double red[kThreads / 64];
  const double pq = sum_partials(part_pq, red);
  const double alpha = (pq != 0.0) ? state[0] / pq : 0.0;   // p = 0 after exact convergence
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    state[2] = alpha;
    state[3] = pq;
  }
  const XT a = (XT)alpha;
  double rz = 0, rr = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    XT ri[3], zi[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      x[3 * i + c] += a * p[3 * i + c];
      ri[c] = r[3 * i + c] - a * q[3 * i + c];
      r[3 * i + c] = ri[c];
    }
    const VT* D = dinv + 9 * i;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      zi[c] = (XT)D[3 * c] * ri[0] + (XT)D[3 * c + 1] * ri[1] + (XT)D[3 * c + 2] * ri[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      z[3 * i + c] = zi[c];
      rz += (double)ri[c] * (double)zi[c];
      rr += (double)ri[c] * (double)ri[c];
    }
  }
  rz = block_sum(rz, red);
  rr = block_sum(rr, red);
  if (threadIdx.x == 0) {
    part_out[blockIdx.x] = rz;
    part_out[kGrid + blockIdx.x] = rr;
  }
}

// beta = sum(part_rz) / state[0]; p = z + beta p; block 0 records
// state[1] = rz_new (moved to state[0] by the next SpMV), state[4] = rr.
template <typename XT>
__global__ void __launch_bounds__(kThreads)
pcg_direction_kernel(int64_t n3, const XT* __restrict__ z, XT* __restrict__ p,
                     const double* __restrict__ part, double* state) {
  __shared__ double red[kThreads / 64];
  const double rz_new = sum_partials(part, red);
  const double beta = (state[0] != 0.0) ? rz_new / state[0] : 0.0;
  if (blockIdx.x == 0) {
    const double rr = sum_partials(part + kGrid, red);
    if (threadIdx.x == 0) {
      state[1] = rz_new;
      state[4] = rr;
      state[6] = beta;
    }
  }
  const XT bt = (XT)beta;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n3; t += (int64_t)gridDim.x * blockDim.x)
    p[t] = z[t] + bt * p[t];
}

// r = b - y; z = Dinv r; partials r.z, r.r (the bb slot is left at 0).
template <typename VT>
__global__ void __launch_bounds__(kThreads)
replace_kernel(int64_t N, const double* __restrict__ b, const double* __restrict__ y,
               const VT* __restrict__ dinv, double* __restrict__ r, double* __restrict__ z,
               double* __restrict__ p, double* __restrict__ part) {
  __shared__ double red[kThreads / 64];
  double rz = 0, rr = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    double ri[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) ri[a] = b[3 * i + a] - y[3 * i + a];
    const VT* D = dinv + 9 * i;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double za = (double)D[3 * a] * ri[0] + (double)D[3 * a + 1] * ri[1] + (double)D[3 * a + 2] * ri[2];
      r[3 * i + a] = ri[a];
      z[3 * i + a] = za;
      if (p) p[3 * i + a] = za;
      rz += ri[a] * za;
      rr += ri[a] * ri[a];
    }
  }
  rz = block_sum(rz, red);
  rr = block_sum(rr, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = rz;
    part[kGrid + blockIdx.x] = rr;
    part[2 * kGrid + blockIdx.x] = 0.0;
  }
}

int check_solve_mesh(const fem_mesh* m) {
  FEM_REQUIRE(m != nullptr, FEM_E_INVALID, "mesh is NULL");
  FEM_REQUIRE(m->num_node >= 0 && m->nnzb >= 0, FEM_E_INVALID, "negative size");
  FEM_REQUIRE(m->num_node == 0 || (m->sm_node && m->co_node), FEM_E_INVALID, "pattern missing");
  FEM_REQUIRE(m->nnzb == 0 || m->re_node, FEM_E_INVALID, "re_node missing");
  return FEM_OK;
}

int64_t grid_for(int64_t items, int per_block) {
  int64_t g = ceil_div(items > 0 ? items : 1, per_block);
  return g < kGrid ? g : kGrid;
}

}  // namespace
}  // namespace fem

using namespace fem;

extern "C" int64_t fem_pcg_work_bytes(int64_t num_node) {
  // six kGrid-sized partial areas + one partial per SpMV work-group
  return 8 * ((int64_t)kGrid * 6 + ceil_div(num_node, kRowsPerBlock)) + 256;
}

extern "C" int fem_bsr_spmv(const fem_mesh* mesh, const void* vals, int dtype, const void* x, void* y,
                            void* stream) {
  int rc = check_solve_mesh(mesh);
  if (rc) return rc;
  if (mesh->num_node == 0) return FEM_OK;
  FEM_REQUIRE(vals && x && y, FEM_E_INVALID, "null pointer");
  hipStream_t st = as_stream(stream);
  const int64_t nb = ceil_div(mesh->num_node, kRowsPerBlock);
  if (dtype == FEM_DTYPE_F64)
    spmv_kernel<double, double, false><<<nb, kThreads, 0, st>>>(*mesh, (const double*)vals, (const double*)x,
                                                                (double*)y, nullptr, nullptr);
  else if (dtype == FEM_DTYPE_F32)
    spmv_kernel<float, float, false><<<nb, kThreads, 0, st>>>(*mesh, (const float*)vals, (const float*)x,
                                                              (float*)y, nullptr, nullptr);
  else if (dtype == FEM_DTYPE_F32F64)
    spmv_kernel<float, double, false><<<nb, kThreads, 0, st>>>(*mesh, (const float*)vals, (const double*)x,
                                                               (double*)y, nullptr, nullptr);
  else
    FEM_REQUIRE(false, FEM_E_INVALID, "bad dtype");
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_bsr_spmm(const fem_mesh* mesh, const double* vals, const double* X, double* Y,
                            int64_t nrhs, void* stream) {
  int rc = check_solve_mesh(mesh);
  if (rc) return rc;
  FEM_REQUIRE(nrhs > 0 && nrhs <= 1024, FEM_E_INVALID, "nrhs must be in [1, 1024]");
  if (mesh->num_node == 0) return FEM_OK;
  FEM_REQUIRE(vals && X && Y, FEM_E_INVALID, "null pointer");
  const bool al16 = reinterpret_cast<uintptr_t>(X) % 16 == 0 && reinterpret_cast<uintptr_t>(Y) % 16 == 0;
  if (nrhs == 16 && al16) {
    spmm16_kernel<<<ceil_div(mesh->num_node * kS16Team, 256), 256, 0, as_stream(stream)>>>(*mesh, vals, X, Y);
    FEM_LAUNCH_CHECK();
    return FEM_OK;
  }
  if (nrhs >= 8 && nrhs % 2 == 0 && al16) {
    const int64_t T2 = mesh->num_node * (nrhs / 2);
    spmm2_kernel<<<ceil_div(T2, 256), 256, 0, as_stream(stream)>>>(*mesh, vals, X, Y, (int)nrhs);
    FEM_LAUNCH_CHECK();
    return FEM_OK;
  }
  const int64_t T = mesh->num_node * nrhs;
  spmm_kernel<<<ceil_div(T, 256), 256, 0, as_stream(stream)>>>(*mesh, vals, X, Y, (int)nrhs);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_mass_spmm(const fem_mesh* mesh, const double* m_blocks, const double* X, double* Y,
                             int64_t nrhs, void* stream) {
  int rc = check_solve_mesh(mesh);
  if (rc) return rc;
  FEM_REQUIRE(nrhs > 0 && nrhs <= 1024, FEM_E_INVALID, "nrhs must be in [1, 1024]");
  if (mesh->num_node == 0) return FEM_OK;
  FEM_REQUIRE(m_blocks && X && Y, FEM_E_INVALID, "null pointer");
  const int64_t T = mesh->num_node * nrhs;
  mass_spmm_kernel<<<ceil_div(T, 256), 256, 0, as_stream(stream)>>>(*mesh, m_blocks, X, Y, (int)nrhs);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_bj_setup(const fem_mesh* mesh, const void* vals, int dtype, void* dinv, int32_t* flag,
                            void* stream) {
  int rc = check_solve_mesh(mesh);
  if (rc) return rc;
  if (mesh->num_node == 0) return FEM_OK;
  FEM_REQUIRE(vals && dinv, FEM_E_INVALID, "null pointer");
  hipStream_t st = as_stream(stream);
  const int64_t nb = ceil_div(mesh->num_node, 256);
  if (dtype == FEM_DTYPE_F64)
    bj_setup_kernel<double><<<nb, 256, 0, st>>>(*mesh, (const double*)vals, (double*)dinv, flag);
  else if (dtype == FEM_DTYPE_F32)
    bj_setup_kernel<float><<<nb, 256, 0, st>>>(*mesh, (const float*)vals, (float*)dinv, flag);
  else
    FEM_REQUIRE(false, FEM_E_INVALID, "bad dtype");
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

template <typename VT, typename XT>
static int pcg_start_t(const fem_mesh* mesh, const VT* vals, const VT* dinv, const XT* b, XT* x, XT* r, XT* z,
                       XT* p, double* state, double* part, int with_bb, hipStream_t st) {
  const int64_t N = mesh->num_node;
  // z := A x as a temporary, then r = b - z etc.
  spmv_kernel<VT, XT, false><<<ceil_div(N, kRowsPerBlock), kThreads, 0, st>>>(*mesh, vals, x, z, nullptr,
                                                                              nullptr);
  pcg_start_kernel<VT, XT><<<kGrid, kThreads, 0, st>>>(N, b, z, dinv, r, z, p, part);
  pcg_start_finalize<<<1, kThreads, 0, st>>>(part, state, with_bb);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

template <typename VT, typename XT>
static int pcg_iterate_t(const fem_mesh* mesh, const VT* vals, const VT* dinv, XT* x, XT* r, XT* z, XT* p,
                         XT* q, double* state, double* part, hipStream_t st) {
  const int64_t N = mesh->num_node;
  // the SpMV runs on the full row grid (one partial p.q per work-group, into the
  // scratch after the kGrid-sized areas); the fold sums them, in order, into kGrid partials
  const int64_t nblk = ceil_div(N, kRowsPerBlock);
  double* rows_part = part + 6 * kGrid;
  spmv_kernel<VT, XT, true><<<nblk, kThreads, 0, st>>>(*mesh, vals, p, q, rows_part, state);
  fold_partials_kernel<<<kGrid, kThreads, 0, st>>>(rows_part, nblk, part);
  pcg_update_kernel<VT, XT><<<kGrid, kThreads, 0, st>>>(N, x, r, z, p, q, dinv, part,
                                                                       part + kGrid, state);
  pcg_direction_kernel<XT><<<grid_for(3 * N, kThreads), kThreads, 0, st>>>(3 * N, z, p, part + kGrid, state);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

#define FEM_PCG_DISPATCH(FN, ...)                                                                  \
  if (dtype == FEM_DTYPE_F64) return FN<double, double>(__VA_ARGS__);                             \
  if (dtype == FEM_DTYPE_F32) return FN<float, float>(__VA_ARGS__);                               \
  if (dtype == FEM_DTYPE_F32F64) return FN<float, double>(__VA_ARGS__);                           \
  FEM_REQUIRE(false, FEM_E_INVALID, "bad dtype %d", dtype);

extern "C" int fem_pcg_start(const fem_mesh* mesh, const void* vals, const void* dinv, int dtype,
                             const void* b, void* x, void* r, void* z, void* p, double* state,
                             void* work, int64_t work_bytes, void* stream) {
  int rc = check_solve_mesh(mesh);
  if (rc) return rc;
  if (mesh->num_node == 0) return FEM_OK;
  FEM_REQUIRE(vals && dinv && b && x && r && z && p && state && work, FEM_E_INVALID, "null pointer");
  FEM_REQUIRE(work_bytes >= fem_pcg_work_bytes(mesh->num_node), FEM_E_INVALID, "work too small");
  hipStream_t st = as_stream(stream);
  double* part = reinterpret_cast<double*>(work);
  if (dtype == FEM_DTYPE_F64)
    return pcg_start_t<double, double>(mesh, (const double*)vals, (const double*)dinv, (const double*)b,
                                       (double*)x, (double*)r, (double*)z, (double*)p, state, part, 1, st);
  if (dtype == FEM_DTYPE_F32)
    return pcg_start_t<float, float>(mesh, (const float*)vals, (const float*)dinv, (const float*)b, (float*)x,
                                     (float*)r, (float*)z, (float*)p, state, part, 1, st);
  if (dtype == FEM_DTYPE_F32F64)
    return pcg_start_t<float, double>(mesh, (const float*)vals, (const float*)dinv, (const double*)b,
                                      (double*)x, (double*)r, (double*)z, (double*)p, state, part, 1, st);
  FEM_REQUIRE(false, FEM_E_INVALID, "bad dtype %d", dtype);
}

extern "C" int fem_pcg_iterate(const fem_mesh* mesh, const void* vals, const void* dinv, int dtype,
                               void* x, void* r, void* z, void* p, void* q, double* state, void* work,
                               int64_t work_bytes, void* stream) {
  int rc = check_solve_mesh(mesh);
  if (rc) return rc;
  if (mesh->num_node == 0) return FEM_OK;
  FEM_REQUIRE(vals && dinv && x && r && z && p && q && state && work, FEM_E_INVALID, "null pointer");
  FEM_REQUIRE(work_bytes >= fem_pcg_work_bytes(mesh->num_node), FEM_E_INVALID, "work too small");
  hipStream_t st = as_stream(stream);
  double* part = reinterpret_cast<double*>(work);
  if (dtype == FEM_DTYPE_F64)
    return pcg_iterate_t<double, double>(mesh, (const double*)vals, (const double*)dinv, (double*)x,
                                         (double*)r, (double*)z, (double*)p, (double*)q, state, part, st);
  if (dtype == FEM_DTYPE_F32)
    return pcg_iterate_t<float, float>(mesh, (const float*)vals, (const float*)dinv, (float*)x, (float*)r,
                                       (float*)z, (float*)p, (float*)q, state, part, st);
  if (dtype == FEM_DTYPE_F32F64)
    return pcg_iterate_t<float, double>(mesh, (const float*)vals, (const float*)dinv, (double*)x, (double*)r,
                                        (double*)z, (double*)p, (double*)q, state, part, st);
  FEM_REQUIRE(false, FEM_E_INVALID, "bad dtype %d", dtype);
}

/* Residual replacement for the mixed form: r = b - y (y = A64 x computed by
 * the caller), z = Dinv r, state[1] = r.z (becomes rz at the next SpMV),
 * state[4] = r.r; with p_restart non-NULL also p = z (CG restart). */
extern "C" int fem_pcg_replace_residual(const fem_mesh* mesh, const void* dinv, int dtype, const void* b,
                                        const void* y, void* r, void* z, void* p_restart, double* state,
                                        void* work, int64_t work_bytes, void* stream) {
  int rc = check_solve_mesh(mesh);
  if (rc) return rc;
  if (mesh->num_node == 0) return FEM_OK;
  FEM_REQUIRE(dtype == FEM_DTYPE_F32F64 || dtype == FEM_DTYPE_F64, FEM_E_INVALID, "bad dtype");
  FEM_REQUIRE(dinv && b && y && r && z && state && work, FEM_E_INVALID, "null pointer");
  FEM_REQUIRE(work_bytes >= fem_pcg_work_bytes(mesh->num_node), FEM_E_INVALID, "work too small");
  hipStream_t st = as_stream(stream);
  const int64_t N = mesh->num_node;
  double* part = reinterpret_cast<double*>(work) + 3 * kGrid;
  // r = b - y; z = Dinv r (p untouched: the start kernel writes p = z, so use a scratch p = r)
  if (dtype == FEM_DTYPE_F32F64)
    replace_kernel<float><<<kGrid, kThreads, 0, st>>>(N, (const double*)b, (const double*)y,
                                                                     (const float*)dinv, (double*)r, (double*)z,
                                                                     (double*)p_restart, part);
  else
    replace_kernel<double><<<kGrid, kThreads, 0, st>>>(N, (const double*)b, (const double*)y,
                                                                      (const double*)dinv, (double*)r,
                                                                      (double*)z, (double*)p_restart, part);
  pcg_start_finalize<<<1, kThreads, 0, st>>>(part, state, 0);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}
