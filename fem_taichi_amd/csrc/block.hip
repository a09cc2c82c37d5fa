// Multi-right-hand-side kernels for block Lanczos (SOLVE_eigen.py:45-89):
// block-Jacobi PCG on nb columns at once (the shift-invert solves
// K^-1 (M u) of :65), with per-column scalars kept on the device.
//
// Blocks of vectors are node-interleaved with columns fastest:
// X[(3*i + p) * nb + c].  nb is a power of two (the caller pads with zero
// columns), so a block-level tree reduction over threads keeps the column
// index (thread id mod nb) and sums every column in a fixed order.  As in the
// single-vector PCG, each of kGrid blocks writes per-column partials and the
// consuming kernel re-sums them in index order: no atomics, bit-reproducible.
#include <mutex>

#include "common.hpp"

namespace fem {
namespace {

constexpr int kT = 256;
constexpr int kGridB = 512;
// Threads per work-group of the kernels that stream the n x nb blocks: with
// the fixed grid of kGridB work-groups, 1024 threads fill a CU.
constexpr int kTS = 1024;

// Per-column sum over the block of NT threads: s[tid] holds values whose
// column is tid % nb; result for column c < nb ends in s[c].
template <int NT = kT>
__device__ void col_reduce(double* s, int nb) {
  for (int w = NT / 2; w >= nb; w >>= 1) {
    __syncthreads();
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
  }
  __syncthreads();
}

// out[c] = sum over blocks of part[b * nb + c] for every column c < nb, in
// the same order in every block (out is shared memory).
template <int NT = kT>
__device__ void col_partials(const double* __restrict__ part, int nb, double* s, double* out) {
  // each thread sums blocks b = tid / nb, tid / nb + NT / nb, ... for column tid % nb
  const int cc = threadIdx.x % nb;
  double v = 0.0;
  for (int b = threadIdx.x / nb; b < kGridB; b += NT / nb) v += part[b * nb + cc];
  s[threadIdx.x] = v;
  col_reduce<NT>(s, nb);
  if ((int)threadIdx.x < nb) out[threadIdx.x] = s[threadIdx.x];
  __syncthreads();
}

// Y = A X with per-column partial X.Y (A: BSR 3x3 f64).  Thread = (row i, c).
// kTS threads per work-group on the fixed grid of kGridB: the gathers of X
// (3 nb values per block) need the memory parallelism of a full CU.
__global__ void __launch_bounds__(kTS)
mspmm_dot_kernel(fem_mesh m, const double* __restrict__ vals, const double* __restrict__ X,
                 double* __restrict__ Y, int nb, double* __restrict__ part, double* state) {
  __shared__ double s[kTS];
  const int c = threadIdx.x % nb;
  double d = 0.0;
  const int64_t total = m.num_node * nb;
  for (int64_t t = (int64_t)blockIdx.x * kTS + threadIdx.x; t < total; t += (int64_t)gridDim.x * kTS) {
    const int64_t i = t / nb;
    const int64_t sm = m.sm_node[i];
    const int co = m.co_node[i];
    double a0 = 0, a1 = 0, a2 = 0;
    for (int b = 0; b < co; ++b) {
      const double* v = vals + 9 * (sm + b);
      const int64_t col = m.re_node[sm + b];
      const double x0 = X[(3 * col) * nb + c], x1 = X[(3 * col + 1) * nb + c], x2 = X[(3 * col + 2) * nb + c];
      a0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
      a1 += v[3] * x0 + v[4] * x1 + v[5] * x2;
      a2 += v[6] * x0 + v[7] * x1 + v[8] * x2;
    }
    Y[(3 * i) * nb + c] = a0;
    Y[(3 * i + 1) * nb + c] = a1;
    Y[(3 * i + 2) * nb + c] = a2;
    d += a0 * X[(3 * i) * nb + c] + a1 * X[(3 * i + 1) * nb + c] + a2 * X[(3 * i + 2) * nb + c];
  }
  if (part) {
    s[threadIdx.x] = d;
    col_reduce<kTS>(s, nb);
    if ((int)threadIdx.x < nb) part[blockIdx.x * nb + threadIdx.x] = s[threadIdx.x];
    // move rz_next (row 1) into rz (row 0) for the update kernel
    if (blockIdx.x == 0 && (int)threadIdx.x < nb) state[threadIdx.x] = state[nb + threadIdx.x];
  }
}

// R = B - Q; Z = Dinv R; P = Z; partials rz, rr, bb per column.
__global__ void __launch_bounds__(kT)
mpcg_start_kernel(int64_t N, int nb, const double* __restrict__ B, const double* Q,
                  const double* __restrict__ dinv, double* R, double* Z, double* P,
                  double* __restrict__ part) {
  __shared__ double s[kT];
  const int c = threadIdx.x % nb;
  double rz = 0, rr = 0, bb = 0;
  const int64_t total = N * nb;
  for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < total; t += (int64_t)gridDim.x * kT) {
    const int64_t i = t / nb;
    double r[3], z[3];
    for (int a = 0; a < 3; ++a) r[a] = B[(3 * i + a) * nb + c] - Q[(3 * i + a) * nb + c];
    const double* D = dinv + 9 * i;
    for (int a = 0; a < 3; ++a) z[a] = D[3 * a] * r[0] + D[3 * a + 1] * r[1] + D[3 * a + 2] * r[2];
    for (int a = 0; a < 3; ++a) {
      R[(3 * i + a) * nb + c] = r[a];
      Z[(3 * i + a) * nb + c] = z[a];
      P[(3 * i + a) * nb + c] = z[a];
      rz += r[a] * z[a];
      rr += r[a] * r[a];
      const double bv = B[(3 * i + a) * nb + c];
      bb += bv * bv;
    }
  }
  double vals3[3] = {rz, rr, bb};
  for (int q = 0; q < 3; ++q) {
    s[threadIdx.x] = vals3[q];
    col_reduce(s, nb);
    if ((int)threadIdx.x < nb) part[(q * kGridB + blockIdx.x) * nb + threadIdx.x] = s[threadIdx.x];
    __syncthreads();
  }
}

// state rows: 0 rz, 1 rz_next, 2 alpha, 3 pq, 4 rr, 5 bb, 6 beta, 7 spare.
__global__ void mpcg_start_finalize(const double* __restrict__ part, int nb, double* state) {
  __shared__ double s[kT];
  __shared__ double o[3][64];
  col_partials(part, nb, s, o[0]);
  col_partials(part + kGridB * nb, nb, s, o[1]);
  col_partials(part + 2 * kGridB * nb, nb, s, o[2]);
  if ((int)threadIdx.x < nb) {
    state[1 * nb + threadIdx.x] = o[0][threadIdx.x];
    state[4 * nb + threadIdx.x] = o[1][threadIdx.x];
    state[5 * nb + threadIdx.x] = o[2][threadIdx.x];
  }
}

__global__ void __launch_bounds__(kTS)
mpcg_update_kernel(int64_t N, int nb, double* __restrict__ X, double* __restrict__ R, double* __restrict__ Z,
                   const double* __restrict__ P, const double* __restrict__ Q, const double* __restrict__ dinv,
                   const double* __restrict__ part_pq, double* __restrict__ part_out, double* state) {
  __shared__ double s[kTS];
  __shared__ double pq_s[64];
  __shared__ double alpha_s[64];
  col_partials<kTS>(part_pq, nb, s, pq_s);
  if ((int)threadIdx.x < nb) {
    const int c = threadIdx.x;
    const double pq = pq_s[c];
    const double a = (pq != 0.0) ? state[c] / pq : 0.0;
    alpha_s[c] = a;
    if (blockIdx.x == 0) {
      state[2 * nb + c] = a;
      state[3 * nb + c] = pq;
    }
  }
  __syncthreads();
  const int c = threadIdx.x % nb;
  const double a = alpha_s[c];
  double rz = 0, rr = 0;
  const int64_t total = N * nb;
  for (int64_t t = (int64_t)blockIdx.x * kTS + threadIdx.x; t < total; t += (int64_t)gridDim.x * kTS) {
    const int64_t i = t / nb;
    double r[3], z[3];
    for (int q = 0; q < 3; ++q) {
      const int64_t o = (3 * i + q) * nb + c;
      X[o] += a * P[o];
      r[q] = R[o] - a * Q[o];
      R[o] = r[q];
    }
    const double* D = dinv + 9 * i;
    for (int q = 0; q < 3; ++q) {
      z[q] = D[3 * q] * r[0] + D[3 * q + 1] * r[1] + D[3 * q + 2] * r[2];
      Z[(3 * i + q) * nb + c] = z[q];
      rz += r[q] * z[q];
      rr += r[q] * r[q];
    }
  }
  s[threadIdx.x] = rz;
  col_reduce<kTS>(s, nb);
  if ((int)threadIdx.x < nb) part_out[blockIdx.x * nb + threadIdx.x] = s[threadIdx.x];
  __syncthreads();
  s[threadIdx.x] = rr;
  col_reduce<kTS>(s, nb);
  if ((int)threadIdx.x < nb) part_out[(kGridB + blockIdx.x) * nb + threadIdx.x] = s[threadIdx.x];
}

__global__ void __launch_bounds__(kTS)
mpcg_direction_kernel(int64_t N, int nb, const double* __restrict__ Z, double* __restrict__ P,
                      const double* __restrict__ part, double* state) {
  __shared__ double s[kTS];
  __shared__ double rz_s[64];
  __shared__ double rr_s[64];
  __shared__ double beta_s[64];
  col_partials<kTS>(part, nb, s, rz_s);
  if (blockIdx.x == 0) col_partials<kTS>(part + kGridB * nb, nb, s, rr_s);
  if ((int)threadIdx.x < nb) {
    const int c = threadIdx.x;
    const double rz = state[c];
    const double b = (rz != 0.0) ? rz_s[c] / rz : 0.0;
    beta_s[c] = b;
    if (blockIdx.x == 0) {
      state[1 * nb + c] = rz_s[c];
      state[4 * nb + c] = rr_s[c];
      state[6 * nb + c] = b;
    }
  }
  __syncthreads();
  const int c = threadIdx.x % nb;
  const double b = beta_s[c];
  const int64_t total = 3 * N * nb;
  for (int64_t t = (int64_t)blockIdx.x * kTS + threadIdx.x; t < total; t += (int64_t)gridDim.x * kTS)
    P[t] = Z[t] + b * P[t];
}

// C (p x q) = A^T B for tall A (n x p, row stride lda) and B (n x q, row
// stride ldb): the inner products of block Lanczos (SOLVE_eigen.py:56, 67,
// 72-73, 76), whose reduction runs over the n = 3N rows.  Work-group g sums
// its row chunk into a partial p x q block.  Tiles of R rows of A and B are
// staged in LDS (coalesced, rows padded to whole 4-column groups, R from a
// 48 KB budget); each thread owns a 4 x 4 block of outputs (4 values of A
// and 4 of B per row feed 16 fused multiply-adds) and, when the p q / 16
// output blocks leave threads over, the threads split into G row groups
// (rows r = group (mod G)) that are folded in group order at the end.  The
// partials are then summed per output in a fixed tree order, so the result
// is deterministic.
constexpr int kTT = 256;
constexpr int kTnLds = 6144;          // doubles of LDS for one A/B row tile
constexpr int kGridTN = 1024;

// Rows rb.. of a tall matrix (w used columns, row stride ld) into LDS rows of
// W doubles (zero padded), total = rows * W entries: every load of a batch is
// issued before the batch's LDS stores, so a thread has kTnBatch loads in
// flight instead of one.
constexpr int kTnBatch = 8;
__device__ __forceinline__ void tn_stage(double* __restrict__ dst, const double* __restrict__ src, int64_t ld, int w,
                                         int W, int64_t rb, int total, int tid) {
  for (int e0 = tid; e0 < total; e0 += kTnBatch * kTT) {
    double v[kTnBatch];
#pragma unroll
    for (int u = 0; u < kTnBatch; ++u) {
      const int e = e0 + u * kTT;
      const int r = e / W, c = e - r * W;
      v[u] = (e < total && c < w) ? src[(rb + r) * ld + c] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kTnBatch; ++u) {
      const int e = e0 + u * kTT;
      if (e < total) dst[e] = v[u];
    }
  }
}

__global__ void __launch_bounds__(kTT)
gemm_tn_partial_kernel(const double* __restrict__ A, int64_t lda, int64_t n, int p, const double* __restrict__ B,
                       int64_t ldb, int q, int R, double* __restrict__ part) {
  extern __shared__ __align__(16) double sm[];
  const int pt = (p + 3) / 4, qt = (q + 3) / 4, ntile = pt * qt;
  const int PA = 4 * pt, QB = 4 * qt;
  const int G = ntile <= kTT ? kTT / ntile : 1;
  double* sA = sm;                      // R x PA
  double* sB = sm + R * PA;             // R x QB
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = (int64_t)blockIdx.x * chunk;
  const int64_t r1 = r0 + chunk < n ? r0 + chunk : n;
  const int tid = threadIdx.x;
  const int grp = tid / ntile;          // >= G: idle (or, when ntile > 256, a second pass below)
  double acc[4][4];
  for (int t0 = 0; t0 < ntile; t0 += kTT) {         // more than 256 output blocks: passes over the rows
    const int tile = t0 + (G > 1 ? tid % ntile : tid);
    const bool on = G > 1 ? grp < G : tile < ntile;
    const int ti = on ? tile / qt : 0, tj = on ? tile % qt : 0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    for (int64_t rb = r0; rb < r1; rb += R) {
      const int nr = (int)(r1 - rb < R ? r1 - rb : R);
      __syncthreads();
      tn_stage(sA, A, lda, p, PA, rb, nr * PA, tid);
      tn_stage(sB, B, ldb, q, QB, rb, nr * QB, tid);
      __syncthreads();
      if (on) {
        for (int r = G > 1 ? grp : 0; r < nr; r += G > 1 ? G : 1) {
          const double2* pa = reinterpret_cast<const double2*>(sA + r * PA + 4 * ti);
          const double2* pb = reinterpret_cast<const double2*>(sB + r * QB + 4 * tj);
          const double2 a01 = pa[0], a23 = pa[1], b01 = pb[0], b23 = pb[1];
          const double av[4] = {a01.x, a01.y, a23.x, a23.y}, bv[4] = {b01.x, b01.y, b23.x, b23.y};
#pragma unroll
          for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = fma(av[a], bv[b], acc[a][b]);
        }
      }
    }
    // fold the row groups in order, then store the block's partial outputs
    __syncthreads();
    if (on && G > 1) {
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) sm[(grp * ntile + tile) * 16 + 4 * a + b] = acc[a][b];
    }
    __syncthreads();
    if (G > 1) {
      for (int o = tid; o < ntile * 16; o += kTT) {
        double v = 0.0;
        for (int g = 0; g < G; ++g) v += sm[g * ntile * 16 + o];
        const int tl = o / 16, ab = o % 16;
        const int i = (tl / qt) * 4 + ab / 4, jj = (tl % qt) * 4 + ab % 4;
        if (i < p && jj < q) part[(int64_t)blockIdx.x * p * q + i * q + jj] = v;
      }
    } else if (on) {
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int i = 4 * ti + a, jj = 4 * tj + b;
          if (i < p && jj < q) part[(int64_t)blockIdx.x * p * q + i * q + jj] = acc[a][b];
        }
    }
  }
}

// One work-group per output: thread t sums partials t, t + 256, ... in order,
// then a fixed LDS tree.
__global__ void __launch_bounds__(kTT)
gemm_tn_sum_kernel(const double* __restrict__ part, int pq, int nblk, double* __restrict__ C) {
  __shared__ double red[kTT];
  const int o = blockIdx.x;
  double s = 0.0;
  for (int g = threadIdx.x; g < nblk; g += kTT) s += part[(int64_t)g * pq + o];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kTT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) C[o] = red[0];
}

// ---- flexible CG, FCG(1), vector steps on nb column-interleaved vectors
// (entry (i, c) at i * nb + c): the Krylov updates around the multigrid
// preconditioner, fused so that each step reads each vector once.  Work-group
// b owns rows [b * chunk, (b + 1) * chunk); thread t takes column t % nb and
// rows t / nb, t / nb + G, ... (G = 256 / nb row subgroups), so a wave reads
// contiguous memory.  Each dot product leaves one partial per (work-group,
// column), folded over subgroups in order, and a one-work-group kernel folds
// the partials in work-group order: deterministic, no atomics.
constexpr int kFT = 256;
constexpr int kFcgGrid = 4096;   // 4,096 x 256 threads: enough loads in flight to stream HBM
// per-column scalars after the partials (doubles)
enum { kScBeta = 0, kScAlpha = 1, kScPqPrev = 2, kScCount = 4 };

struct FcgMap {
  int64_t r0, r1;
  int c, s, G;
  bool on;
};

__device__ __forceinline__ FcgMap fcg_map(int64_t n, int nb) {
  FcgMap f;
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  f.r0 = (int64_t)blockIdx.x * chunk;
  f.r1 = f.r0 + chunk < n ? f.r0 + chunk : n;
  f.G = kFT / nb;
  f.c = (int)threadIdx.x % nb;
  f.s = (int)threadIdx.x / nb;
  f.on = f.s < f.G;
  return f;
}

// Sum of v over the row subgroups of each column, in subgroup order, written
// to out[c] by thread c (c < nb).
__device__ __forceinline__ void fcg_block_fold(double v, const FcgMap& f, int nb, double* red, double* out) {
  __syncthreads();
  red[threadIdx.x] = f.on ? v : 0.0;
  __syncthreads();
  if ((int)threadIdx.x < nb) {
    double s = 0.0;
    for (int g = 0; g < f.G; ++g) s += red[g * nb + threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// partial[b][c] = sum over the work-group's rows of mask_i * A_ic * B_ic (A FP32 or FP64)
template <typename TA>
__global__ void __launch_bounds__(kFT)
fcg_dot_kernel(int64_t n, int nb, const TA* __restrict__ A, const double* __restrict__ B,
               const double* __restrict__ mask, double* __restrict__ part) {
  __shared__ double red[kFT];
  const FcgMap f = fcg_map(n, nb);
  double acc = 0.0;
  if (f.on)
#pragma unroll 4
    for (int64_t i = f.r0 + f.s; i < f.r1; i += f.G) {
      const double a = (double)A[i * nb + f.c] * B[i * nb + f.c];
      acc += mask ? a * mask[i] : a;
    }
  fcg_block_fold(acc, f, nb, red, part + (int64_t)blockIdx.x * nb);
}

// Two partials at once: sum P*Q and sum P*R.
__global__ void __launch_bounds__(kFT)
fcg_dot2_kernel(int64_t n, int nb, const double* __restrict__ P, const double* __restrict__ Q,
                const double* __restrict__ R, double* __restrict__ part_pq, double* __restrict__ part_pr) {
  __shared__ double red[kFT];
  const FcgMap f = fcg_map(n, nb);
  double pq = 0.0, pr = 0.0;
  if (f.on)
#pragma unroll 4
    for (int64_t i = f.r0 + f.s; i < f.r1; i += f.G) {
      const int64_t e = i * nb + f.c;
      const double p = P[e];
      pq += p * Q[e];
      pr += p * R[e];
    }
  fcg_block_fold(pq, f, nb, red, part_pq + (int64_t)blockIdx.x * nb);
  fcg_block_fold(pr, f, nb, red, part_pr + (int64_t)blockIdx.x * nb);
}

// Column totals of nblk partials by one 1024-thread work-group: thread
// (subset s, column c) sums partials s, s + G, ... (G = min(1024 / nb, 256)
// subsets, up to eight loads in flight), then thread c sums the subsets in
// order.
constexpr int kFoldT = 1024;

__device__ __forceinline__ double fcg_fold_partials(const double* __restrict__ part, int nblk, int nb, double* red,
                                                    int* col) {
  const int G = kFoldT / nb < 256 ? kFoldT / nb : 256;
  const int c = (int)threadIdx.x % nb, s = (int)threadIdx.x / nb;
  double v = 0.0;
  if (s < G) {
#pragma unroll 8
    for (int g = s; g < nblk; g += G) v += part[(int64_t)g * nb + c];
  }
  __syncthreads();
  red[threadIdx.x] = s < G ? v : 0.0;
  __syncthreads();
  double tot = 0.0;
  if ((int)threadIdx.x < nb)
    for (int g = 0; g < G; ++g) tot += red[g * nb + threadIdx.x];
  *col = (int)threadIdx.x;
  return tot;
}

// beta_c = (sum mask Z Qp)_c / pq_prev_c
__global__ void __launch_bounds__(kFoldT)
fcg_beta_kernel(const double* __restrict__ part, int nblk, int nb, double* __restrict__ sc) {
  __shared__ double red[kFoldT];
  int c;
  const double zq = fcg_fold_partials(part, nblk, nb, red, &c);
  if (c < nb) sc[kScBeta * nb + c] = zq / sc[kScPqPrev * nb + c];
}

// alpha_c = pr / pq on active columns (0 on frozen ones); pq kept for beta
__global__ void __launch_bounds__(kFoldT)
fcg_alpha_kernel(const double* __restrict__ part_pq, const double* __restrict__ part_pr, int nblk, int nb,
                 const int32_t* __restrict__ active, double* __restrict__ sc) {
  __shared__ double red[kFoldT];
  int c;
  const double pq = fcg_fold_partials(part_pq, nblk, nb, red, &c);
  const double pr = fcg_fold_partials(part_pr, nblk, nb, red, &c);
  if (c < nb) {
    const double pq_safe = pq != 0.0 ? pq : 1.0;
    sc[kScAlpha * nb + c] = (active == nullptr || active[c]) ? pr / pq_safe : 0.0;
    sc[kScPqPrev * nb + c] = pq_safe;
  }
}

// alpha_c = pr / pq from partial sets of different counts (pq from the
// structured-grid product's work-groups, pr from the direction pass)
__global__ void __launch_bounds__(kFoldT)
fcg_alpha_split_kernel(const double* __restrict__ part_pq, int nblk_pq, const double* __restrict__ part_pr,
                       int nblk_pr, int nb, const int32_t* __restrict__ active, double* __restrict__ sc) {
  __shared__ double red[kFoldT];
  int c;
  const double pq = fcg_fold_partials(part_pq, nblk_pq, nb, red, &c);
  const double pr = fcg_fold_partials(part_pr, nblk_pr, nb, red, &c);
  if (c < nb) {
    const double pq_safe = pq != 0.0 ? pq : 1.0;
    sc[kScAlpha * nb + c] = (active == nullptr || active[c]) ? pr / pq_safe : 0.0;
    sc[kScPqPrev * nb + c] = pq_safe;
  }
}

__global__ void __launch_bounds__(kFoldT)
fcg_sum_kernel(const double* __restrict__ part, int nblk, int nb, double* __restrict__ out) {
  __shared__ double red[kFoldT];
  int c;
  const double v = fcg_fold_partials(part, nblk, nb, red, &c);
  if (c < nb) out[c] = v;
}

// P = mask * Z - beta * P   (first: P = mask * Z); PR: also the partials
// of sum P * R (the numerator of alpha) of the new P
template <typename TZ, bool PR>
__global__ void __launch_bounds__(kFT)
fcg_dir_kernel(int64_t n, int nb, const TZ* __restrict__ Z, const double* __restrict__ mask,
               const double* __restrict__ sc, int first, double* __restrict__ P, const double* __restrict__ R,
               double* __restrict__ part_pr) {
  const FcgMap f = fcg_map(n, nb);
  double pr = 0.0;
  if (f.on) {
    const double beta = first ? 0.0 : sc[kScBeta * nb + f.c];
#pragma unroll 4
    for (int64_t i = f.r0 + f.s; i < f.r1; i += f.G) {
      const int64_t e = i * nb + f.c;
      double z = (double)Z[e];
      if (mask) z *= mask[i];
      const double p = first ? z : z - beta * P[e];
      P[e] = p;
      if (PR) pr += p * R[e];
    }
  }
  if (PR) {
    __shared__ double red[kFT];
    fcg_block_fold(pr, f, nb, red, part_pr + (int64_t)blockIdx.x * nb);
  }
}

// X += alpha P; R -= alpha * mask * Q; R32 = (float)R; partial sum R^2
__global__ void __launch_bounds__(kFT)
fcg_update_kernel(int64_t n, int nb, const double* __restrict__ P, const double* __restrict__ Q,
                  const double* __restrict__ mask, const double* __restrict__ sc, double* __restrict__ X,
                  double* __restrict__ R, float* __restrict__ R32, double* __restrict__ part) {
  __shared__ double red[kFT];
  const FcgMap f = fcg_map(n, nb);
  double rr = 0.0;
  if (f.on) {
    const double alpha = sc[kScAlpha * nb + f.c];
#pragma unroll 4
    for (int64_t i = f.r0 + f.s; i < f.r1; i += f.G) {
      const int64_t e = i * nb + f.c;
      X[e] += alpha * P[e];
      const double q = mask ? Q[e] * mask[i] : Q[e];
      const double r = R[e] - alpha * q;
      R[e] = r;
      if (R32) R32[e] = (float)r;
      rr += r * r;
    }
  }
  fcg_block_fold(rr, f, nb, red, part + (int64_t)blockIdx.x * nb);
}

// t = (B - Y) * mask_i (Y NULL: t = B * mask_i; mask NULL: all ones);
// partials of sum t^2 and sum B^2 per column; columns with sel[c] != 0 (all
// when sel is NULL) also get R = t and R32 = (float)t.
__global__ void __launch_bounds__(kFT)
fcg_residual_kernel(int64_t n, int nb, const double* __restrict__ B, const double* __restrict__ Y,
                    const double* __restrict__ mask, const int32_t* __restrict__ sel, double* __restrict__ R,
                    float* __restrict__ R32, double* __restrict__ part_t, double* __restrict__ part_b) {
  __shared__ double red[kFT];
  const FcgMap f = fcg_map(n, nb);
  double tt = 0.0, bb = 0.0;
  if (f.on) {
    const bool wr = sel == nullptr || sel[f.c] != 0;
#pragma unroll 4
    for (int64_t i = f.r0 + f.s; i < f.r1; i += f.G) {
      const int64_t e = i * nb + f.c;
      const double b = B[e];
      double t = Y ? b - Y[e] : b;
      if (mask) t *= mask[i];
      tt += t * t;
      bb += b * b;
      if (wr) {
        if (R) R[e] = t;
        if (R32) R32[e] = (float)t;
      }
    }
  }
  fcg_block_fold(tt, f, nb, red, part_t + (int64_t)blockIdx.x * nb);
  fcg_block_fold(bb, f, nb, red, part_b + (int64_t)blockIdx.x * nb);
}

int fcg_grid(int64_t n) { return (int)(n < kFcgGrid ? (n < 1 ? 1 : n) : kFcgGrid); }

int check_nb(int64_t nb) {
  FEM_REQUIRE(nb >= 1 && nb <= 64 && (nb & (nb - 1)) == 0, FEM_E_INVALID,
              "nrhs must be a power of two in [1, 64] (pad with zero columns), got %lld", (long long)nb);
  return FEM_OK;
}

}  // namespace
}  // namespace fem

using namespace fem;

extern "C" int64_t fem_mpcg_work_bytes(int64_t nrhs) { return 8 * (int64_t)kGridB * nrhs * 6 + 256; }

extern "C" int fem_mpcg_start(const fem_mesh* mesh, const double* vals, const double* dinv, int64_t nrhs,
                              const double* B, double* X, double* R, double* Z, double* P, double* state,
                              void* work, int64_t work_bytes, void* stream) {
  FEM_REQUIRE(mesh && vals && dinv && B && X && R && Z && P && state && work, FEM_E_INVALID, "null pointer");
  int rc = check_nb(nrhs);
  if (rc) return rc;
  FEM_REQUIRE(work_bytes >= fem_mpcg_work_bytes(nrhs), FEM_E_INVALID, "work too small");
  if (mesh->num_node == 0) return FEM_OK;
  hipStream_t st = as_stream(stream);
  double* part = reinterpret_cast<double*>(work);
  const int nb = (int)nrhs;
  mspmm_dot_kernel<<<kGridB, kTS, 0, st>>>(*mesh, vals, X, Z, nb, nullptr, nullptr);
  mpcg_start_kernel<<<kGridB, kT, 0, st>>>(mesh->num_node, nb, B, Z, dinv, R, Z, P, part);
  mpcg_start_finalize<<<1, kT, 0, st>>>(part, nb, state);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_mpcg_iterate(const fem_mesh* mesh, const double* vals, const double* dinv, int64_t nrhs,
                                double* X, double* R, double* Z, double* P, double* Q, double* state,
                                void* work, int64_t work_bytes, void* stream) {
  FEM_REQUIRE(mesh && vals && dinv && X && R && Z && P && Q && state && work, FEM_E_INVALID, "null pointer");
  int rc = check_nb(nrhs);
  if (rc) return rc;
  FEM_REQUIRE(work_bytes >= fem_mpcg_work_bytes(nrhs), FEM_E_INVALID, "work too small");
  if (mesh->num_node == 0) return FEM_OK;
  hipStream_t st = as_stream(stream);
  double* part = reinterpret_cast<double*>(work);
  const int nb = (int)nrhs;
  const int64_t N = mesh->num_node;
  mspmm_dot_kernel<<<kGridB, kTS, 0, st>>>(*mesh, vals, P, Q, nb, part, state);
  mpcg_update_kernel<<<kGridB, kTS, 0, st>>>(N, nb, X, R, Z, P, Q, dinv, part, part + kGridB * nb, state);
  mpcg_direction_kernel<<<kGridB, kTS, 0, st>>>(N, nb, Z, P, part + kGridB * nb, state);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int64_t fem_gemm_tn_work_bytes(int64_t p, int64_t q) { return 8 * (int64_t)kGridTN * p * q + 256; }

namespace fem {
namespace {

int gemm_tn_impl(const double* A, int64_t lda, const double* B, int64_t ldb, int64_t n, int64_t p, int64_t q,
                 double* C, double* part, hipStream_t st) {
  const int pi = (int)p, qi = (int)q, pqi = (int)(p * q);
  const int PA = 4 * ((pi + 3) / 4), QB = 4 * ((qi + 3) / 4);
  int R = kTnLds / (PA + QB);
  if (R > 256) R = 256;
  const size_t lds = (size_t)kTnLds * 8;      // also holds the row-group fold (<= 4096 doubles)
  const int nblk = n < 1 ? 1 : (int)(n < kGridTN ? n : kGridTN);
  gemm_tn_partial_kernel<<<nblk, kTT, lds, st>>>(A, lda, n, pi, B, ldb, qi, R, part);
  FEM_LAUNCH_CHECK();
  gemm_tn_sum_kernel<<<pqi, kTT, 0, st>>>(part, pqi, nblk, C);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

int check_gemm_tn(const double* A, int64_t lda, const double* B, int64_t ldb, int64_t n, int64_t p, int64_t q,
                  const double* C) {
  FEM_REQUIRE(A && B && C, FEM_E_INVALID, "null pointer");
  FEM_REQUIRE(n >= 0 && p >= 1 && q >= 1 && p <= 512 && q <= 64 && p * q <= 32 * kTT, FEM_E_INVALID,
              "need 1 <= p <= 512, 1 <= q <= 64, p*q <= %d (got %lld x %lld)", 32 * kTT, (long long)p,
              (long long)q);
  FEM_REQUIRE(lda >= p && ldb >= q, FEM_E_INVALID, "row strides lda %lld < p or ldb %lld < q", (long long)lda,
              (long long)ldb);
  return FEM_OK;
}

// The no-allocation variant's scratch: the partial blocks of at most
// kNoAllocPQ outputs, kGridTN x kNoAllocPQ doubles (16 MiB), one copy per GPU
// for the life of the process.  Its device address is looked up once per
// device (hipGetSymbolAddress with that device current) and kept here; host
// code passes only that address to kernels, never the array's name.
constexpr int kNoAllocPQ = FEM_GEMM_TN_NOALLOC_MAX_PQ;
constexpr int kMaxDevices = 64;
__device__ __align__(256) double scratch[(size_t)kGridTN * kNoAllocPQ];
double* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_for_current_device(double** out) {
  int dev = 0;
  FEM_CHECK_HIP(hipGetDevice(&dev));
  FEM_REQUIRE(dev >= 0 && dev < kMaxDevices, FEM_E_LIMIT, "device %d beyond the %d supported", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    FEM_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[dev] = reinterpret_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return FEM_OK;
}

}  // namespace
}  // namespace fem

extern "C" int fem_gemm_tn_f64(const double* A, int64_t lda, const double* B, int64_t ldb, int64_t n, int64_t p,
                               int64_t q, double* C, void* work, int64_t work_bytes, void* stream) {
  int rc = check_gemm_tn(A, lda, B, ldb, n, p, q, C);
  if (rc) return rc;
  FEM_REQUIRE(work, FEM_E_INVALID, "null pointer");
  FEM_REQUIRE(work_bytes >= fem_gemm_tn_work_bytes(p, q), FEM_E_INVALID, "work too small");
  return gemm_tn_impl(A, lda, B, ldb, n, p, q, C, reinterpret_cast<double*>(work), as_stream(stream));
}

extern "C" int fem_gemm_tn_f64_noalloc(const double* A, int64_t lda, const double* B, int64_t ldb, int64_t n,
                                       int64_t p, int64_t q, double* C, void* stream) {
  int rc = check_gemm_tn(A, lda, B, ldb, n, p, q, C);
  if (rc) return rc;
  FEM_REQUIRE(p * q <= kNoAllocPQ, FEM_E_LIMIT,
              "p*q = %lld exceeds the %d outputs of the static scratch (use fem_gemm_tn_f64 with a workspace)",
              (long long)(p * q), kNoAllocPQ);
  double* part = nullptr;
  rc = scratch_for_current_device(&part);
  if (rc) return rc;
  return gemm_tn_impl(A, lda, B, ldb, n, p, q, C, part, as_stream(stream));
}

// ---- FCG(1) vector steps -------------------------------------------------
// work: three partial arrays (kFcgGrid x nb doubles each), then the
// per-column scalars (beta, alpha, the previous p.q), which carry from
// fem_fcg_update to the next fem_fcg_direction: pass the same work buffer.
extern "C" int64_t fem_fcg_work_bytes(int64_t nb) { return 8 * (3 * (int64_t)kFcgGrid * nb + kScCount * nb) + 256; }

extern "C" int fem_fcg_direction(int64_t n, int64_t nb, const void* Z, int z_dtype, const double* mask,
                                 const double* Qp, double* P, int first, void* work, int64_t work_bytes,
                                 void* stream) {
  FEM_REQUIRE(nb >= 1 && nb <= 64, FEM_E_INVALID, "nb must be in [1, 64] (got %lld)", (long long)nb);
  FEM_REQUIRE(z_dtype == FEM_DTYPE_F32 || z_dtype == FEM_DTYPE_F64, FEM_E_INVALID, "bad z_dtype %d", z_dtype);
  FEM_REQUIRE(n >= 0, FEM_E_INVALID, "negative row count");
  FEM_REQUIRE(work && work_bytes >= fem_fcg_work_bytes(nb), FEM_E_INVALID, "work too small");
  if (n == 0) return FEM_OK;
  FEM_REQUIRE(Z && P && (first || Qp), FEM_E_INVALID, "null pointer");
  hipStream_t st = as_stream(stream);
  double* part = reinterpret_cast<double*>(work);
  double* sc = part + 3 * (int64_t)kFcgGrid * nb;
  const int grid = fcg_grid(n), nbi = (int)nb;
  const bool f32 = z_dtype == FEM_DTYPE_F32;
  const float* Zf = reinterpret_cast<const float*>(Z);
  const double* Zd = reinterpret_cast<const double*>(Z);
  if (!first) {
    if (f32) fcg_dot_kernel<float><<<grid, kFT, 0, st>>>(n, nbi, Zf, Qp, mask, part);
    else fcg_dot_kernel<double><<<grid, kFT, 0, st>>>(n, nbi, Zd, Qp, mask, part);
    FEM_LAUNCH_CHECK();
    fcg_beta_kernel<<<1, kFoldT, 0, st>>>(part, grid, nbi, sc);
    FEM_LAUNCH_CHECK();
  }
  if (f32) fcg_dir_kernel<float, false><<<grid, kFT, 0, st>>>(n, nbi, Zf, mask, sc, first ? 1 : 0, P, nullptr, nullptr);
  else fcg_dir_kernel<double, false><<<grid, kFT, 0, st>>>(n, nbi, Zd, mask, sc, first ? 1 : 0, P, nullptr, nullptr);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_fcg_update(int64_t n, int64_t nb, const double* P, const double* Q, const double* mask,
                              const int32_t* active, double* X, double* R, float* R32, double* rr, void* work,
                              int64_t work_bytes, void* stream) {
  FEM_REQUIRE(nb >= 1 && nb <= 64, FEM_E_INVALID, "nb must be in [1, 64] (got %lld)", (long long)nb);
  FEM_REQUIRE(n >= 0, FEM_E_INVALID, "negative row count");
  FEM_REQUIRE(work && work_bytes >= fem_fcg_work_bytes(nb), FEM_E_INVALID, "work too small");
  FEM_REQUIRE(rr, FEM_E_INVALID, "null rr");
  hipStream_t st = as_stream(stream);
  double* part = reinterpret_cast<double*>(work);
  double* part2 = part + (int64_t)kFcgGrid * nb;
  double* part3 = part2 + (int64_t)kFcgGrid * nb;
  double* sc = part + 3 * (int64_t)kFcgGrid * nb;
  const int grid = fcg_grid(n), nbi = (int)nb;
  if (n > 0) FEM_REQUIRE(P && Q && X && R, FEM_E_INVALID, "null pointer");
  fcg_dot2_kernel<<<grid, kFT, 0, st>>>(n, nbi, P, Q, R, part, part2);
  FEM_LAUNCH_CHECK();
  fcg_alpha_kernel<<<1, kFoldT, 0, st>>>(part, part2, grid, nbi, active, sc);
  FEM_LAUNCH_CHECK();
  fcg_update_kernel<<<grid, kFT, 0, st>>>(n, nbi, P, Q, mask, sc, X, R, R32, part3);
  FEM_LAUNCH_CHECK();
  fcg_sum_kernel<<<1, kFoldT, 0, st>>>(part3, grid, nbi, rr);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_fcg_direction_pr(int64_t n, int64_t nb, const void* Z, int z_dtype, const double* mask,
                                    const double* Qp, const double* R, double* P, int first, void* work,
                                    int64_t work_bytes, void* stream) {
  FEM_REQUIRE(nb >= 1 && nb <= 64, FEM_E_INVALID, "nb must be in [1, 64] (got %lld)", (long long)nb);
  FEM_REQUIRE(z_dtype == FEM_DTYPE_F32 || z_dtype == FEM_DTYPE_F64, FEM_E_INVALID, "bad z_dtype %d", z_dtype);
  FEM_REQUIRE(n >= 0, FEM_E_INVALID, "negative row count");
  FEM_REQUIRE(work && work_bytes >= fem_fcg_work_bytes(nb), FEM_E_INVALID, "work too small");
  hipStream_t st = as_stream(stream);
  double* part = reinterpret_cast<double*>(work);
  double* part2 = part + (int64_t)kFcgGrid * nb;     // sum P R, read by fem_fcg_update_split
  double* sc = part + 3 * (int64_t)kFcgGrid * nb;
  const int grid = fcg_grid(n), nbi = (int)nb;
  if (n > 0) FEM_REQUIRE(Z && P && R && (first || Qp), FEM_E_INVALID, "null pointer");
  const bool f32 = z_dtype == FEM_DTYPE_F32;
  const float* Zf = reinterpret_cast<const float*>(Z);
  const double* Zd = reinterpret_cast<const double*>(Z);
  if (!first && n > 0) {
    if (f32) fcg_dot_kernel<float><<<grid, kFT, 0, st>>>(n, nbi, Zf, Qp, mask, part);
    else fcg_dot_kernel<double><<<grid, kFT, 0, st>>>(n, nbi, Zd, Qp, mask, part);
    FEM_LAUNCH_CHECK();
    fcg_beta_kernel<<<1, kFoldT, 0, st>>>(part, grid, nbi, sc);
    FEM_LAUNCH_CHECK();
  }
  // with no rows the one work-group only writes zero partials
  const int fst = first || n == 0 ? 1 : 0;
  if (f32) fcg_dir_kernel<float, true><<<grid, kFT, 0, st>>>(n, nbi, Zf, mask, sc, fst, P, R, part2);
  else fcg_dir_kernel<double, true><<<grid, kFT, 0, st>>>(n, nbi, Zd, mask, sc, fst, P, R, part2);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_fcg_update_split(int64_t n, int64_t nb, const double* P, const double* Q, const double* mask,
                                    const int32_t* active, const double* pq_part, int64_t pq_blocks, double* X,
                                    double* R, float* R32, double* rr, void* work, int64_t work_bytes, void* stream) {
  FEM_REQUIRE(nb >= 1 && nb <= 64, FEM_E_INVALID, "nb must be in [1, 64] (got %lld)", (long long)nb);
  FEM_REQUIRE(n >= 0, FEM_E_INVALID, "negative row count");
  FEM_REQUIRE(work && work_bytes >= fem_fcg_work_bytes(nb), FEM_E_INVALID, "work too small");
  FEM_REQUIRE(rr && pq_part, FEM_E_INVALID, "null rr or pq_part");
  FEM_REQUIRE(pq_blocks >= 1 && pq_blocks < INT32_MAX, FEM_E_INVALID, "bad pq_blocks %lld", (long long)pq_blocks);
  hipStream_t st = as_stream(stream);
  double* part = reinterpret_cast<double*>(work);
  double* part2 = part + (int64_t)kFcgGrid * nb;
  double* part3 = part2 + (int64_t)kFcgGrid * nb;
  double* sc = part + 3 * (int64_t)kFcgGrid * nb;
  const int grid = fcg_grid(n), nbi = (int)nb;
  if (n > 0) FEM_REQUIRE(P && Q && X && R, FEM_E_INVALID, "null pointer");
  fcg_alpha_split_kernel<<<1, kFoldT, 0, st>>>(pq_part, (int)pq_blocks, part2, grid, nbi, active, sc);
  FEM_LAUNCH_CHECK();
  fcg_update_kernel<<<grid, kFT, 0, st>>>(n, nbi, P, Q, mask, sc, X, R, R32, part3);
  FEM_LAUNCH_CHECK();
  fcg_sum_kernel<<<1, kFoldT, 0, st>>>(part3, grid, nbi, rr);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_fcg_residual(int64_t n, int64_t nb, const double* B, const double* Y, const double* mask,
                                const int32_t* sel, double* R, float* R32, double* out, void* work,
                                int64_t work_bytes, void* stream) {
  FEM_REQUIRE(nb >= 1 && nb <= 64, FEM_E_INVALID, "nb must be in [1, 64] (got %lld)", (long long)nb);
  FEM_REQUIRE(n >= 0, FEM_E_INVALID, "negative row count");
  FEM_REQUIRE(work && work_bytes >= fem_fcg_work_bytes(nb), FEM_E_INVALID, "work too small");
  FEM_REQUIRE(out, FEM_E_INVALID, "null out");
  FEM_REQUIRE(n == 0 || B, FEM_E_INVALID, "null B");
  FEM_REQUIRE(!R || (R != B && R != Y), FEM_E_INVALID, "R must not alias B or Y");
  hipStream_t st = as_stream(stream);
  double* part = reinterpret_cast<double*>(work);
  double* part2 = part + (int64_t)kFcgGrid * nb;   // the first two partial arrays; the carried scalars stay
  const int grid = fcg_grid(n), nbi = (int)nb;
  fcg_residual_kernel<<<grid, kFT, 0, st>>>(n, nbi, B, Y, mask, sel, R, R32, part, part2);
  FEM_LAUNCH_CHECK();
  fcg_sum_kernel<<<1, kFoldT, 0, st>>>(part, grid, nbi, out);
  FEM_LAUNCH_CHECK();
  fcg_sum_kernel<<<1, kFoldT, 0, st>>>(part2, grid, nbi, out + nb);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

namespace fem {
namespace {

// ---- block Lanczos updates (SOLVE_eigen.py:55, 60-61, 66-71) ------------
// Y = alpha A C + beta Y, A tall (n x k, row stride lda), C small (k x q).
// A work-group owns 256 rows, one per thread.  The k columns go in chunks of
// kNnChunk: the group stages its rows' chunk in LDS with coalesced loads
// (16 consecutive threads read one row's 128 contiguous bytes) and the
// chunk of C beside it; each thread then accumulates its row's q outputs
// from LDS (C entries are broadcast reads).  Every output is a sum in
// ascending column order, so results are deterministic.
constexpr int kNnT = 256;
constexpr int kNnChunk = 16;

template <int QMAX>
__global__ void __launch_bounds__(kNnT)
gemm_nn_kernel(const double* __restrict__ A, int64_t lda, const double* __restrict__ C, int64_t ldc,
               double* __restrict__ Y, int64_t ldy, int64_t n, int k, int q, double alpha, double beta) {
  __shared__ double sA[kNnT * (kNnChunk + 1)];
  __shared__ double sC[kNnChunk * QMAX];
  const int64_t row0 = (int64_t)blockIdx.x * kNnT;
  const int tid = threadIdx.x;
  double acc[QMAX];
#pragma unroll
  for (int c = 0; c < QMAX; ++c) acc[c] = 0.0;
  for (int j0 = 0; j0 < k; j0 += kNnChunk) {
    const int kc = k - j0 < kNnChunk ? k - j0 : kNnChunk;
    __syncthreads();
    double v[kNnChunk];                           // all of the thread's loads in flight, then the stores
#pragma unroll
    for (int i = 0; i < kNnChunk; ++i) {
      const int e = tid + i * kNnT;               // row e / 16, column e % 16 of the chunk
      const int r = e / kNnChunk, j = e - r * kNnChunk;
      const int64_t row = row0 + r;
      v[i] = (row < n && j < kc) ? A[row * lda + j0 + j] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < kNnChunk; ++i) {
      const int e = tid + i * kNnT;
      sA[(e / kNnChunk) * (kNnChunk + 1) + e % kNnChunk] = v[i];
    }
    for (int e = tid; e < kNnChunk * QMAX; e += kNnT) {
      const int j = e / QMAX, c = e - j * QMAX;
      sC[e] = (j < kc && c < q) ? C[(int64_t)(j0 + j) * ldc + c] : 0.0;
    }
    __syncthreads();
    for (int j = 0; j < kc; ++j) {
      const double a = sA[tid * (kNnChunk + 1) + j];
#pragma unroll
      for (int c = 0; c < QMAX; ++c) acc[c] = fma(a, sC[j * QMAX + c], acc[c]);
    }
  }
  // Y through LDS, 16 columns at a time, so that consecutive lanes read and
  // write consecutive addresses (a thread's own row would make every store
  // instruction touch 64 lines for 8 bytes each)
#pragma unroll
  for (int c0 = 0; c0 < QMAX; c0 += kNnChunk) {
    __syncthreads();
#pragma unroll
    for (int cc = 0; cc < kNnChunk && cc < QMAX; ++cc) sA[tid * (kNnChunk + 1) + cc] = acc[c0 + cc];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kNnChunk; ++i) {
      const int e = tid + i * kNnT;
      const int r = e / kNnChunk, cc = e - r * kNnChunk;
      const int64_t row = row0 + r;
      const int c = c0 + cc;
      if (row < n && c < q && cc < QMAX) {
        const double a = sA[r * (kNnChunk + 1) + cc];
        double* y = Y + row * ldy + c;
        *y = beta == 0.0 ? alpha * a : fma(alpha, a, beta * *y);
      }
    }
  }
}

// Upper Cholesky factor R (A = R^T R) of a small SPD b x b matrix and its
// inverse, in one work-group (SOLVE_eigen.py:52, 75: scipy cholesky, then the
// inverse the next block is scaled by).  Columns are spread over the threads;
// a non-positive pivot makes R and Rinv NaN (the caller checks the result).
constexpr int kCholT = 64;

__global__ void __launch_bounds__(kCholT)
chol_upper_inv_kernel(const double* __restrict__ A, int b, double* __restrict__ R, double* __restrict__ Rinv) {
  __shared__ double sR[64 * 64];
  __shared__ double sX[64 * 64];
  const int c = threadIdx.x;
  for (int e = c; e < b * b; e += kCholT) sR[(e / b) * 64 + e % b] = A[e];
  __syncthreads();
  for (int j = 0; j < b; ++j) {
    // row j of R: pivot first, then the rest of the row from the updated pivot
    if (c == 0) {
      const double d = sR[j * 64 + j];
      sR[j * 64 + j] = d > 0.0 ? sqrt(d) : __longlong_as_double(0x7ff8000000000000LL);
    }
    __syncthreads();
    const double piv = sR[j * 64 + j];
    if (c > j && c < b) sR[j * 64 + c] /= piv;
    __syncthreads();
    // trailing update A_cc' -= R_jc R_jc' for c, c' > j (thread c owns column c)
    if (c > j && c < b) {
      const double rjc = sR[j * 64 + c];
      for (int i = j + 1; i <= c; ++i) sR[i * 64 + c] -= sR[j * 64 + i] * rjc;
    }
    __syncthreads();
  }
  if (c < b) {
    for (int i = 0; i < b; ++i) R[i * b + c] = i <= c ? sR[i * 64 + c] : 0.0;
    // column c of R^-1 by back substitution, R x = e_c (x in column c of sX)
    for (int i = b - 1; i >= 0; --i) {
      double s = i == c ? 1.0 : 0.0;
      for (int j = i + 1; j <= c; ++j) s -= sR[i * 64 + j] * sX[j * 64 + c];
      sX[i * 64 + c] = i <= c ? s / sR[i * 64 + i] : 0.0;
    }
    for (int i = 0; i < b; ++i) Rinv[i * b + c] = sX[i * 64 + c];
  }
}

}  // namespace
}  // namespace fem

extern "C" int fem_gemm_nn_f64(const double* A, int64_t lda, const double* C, int64_t ldc, double* Y, int64_t ldy,
                               int64_t n, int64_t k, int64_t q, double alpha, double beta, void* stream) {
  FEM_REQUIRE(n >= 0 && k >= 0 && q >= 1 && q <= 64 && k <= (1 << 20), FEM_E_INVALID,
              "need n, k >= 0 and 1 <= q <= 64 (got n %lld, k %lld, q %lld)", (long long)n, (long long)k,
              (long long)q);
  FEM_REQUIRE(lda >= k && ldc >= q && ldy >= q, FEM_E_INVALID, "row strides lda %lld, ldc %lld, ldy %lld too small",
              (long long)lda, (long long)ldc, (long long)ldy);
  if (n == 0) return FEM_OK;
  FEM_REQUIRE(Y && (k == 0 || (A && C)), FEM_E_INVALID, "null pointer");
  hipStream_t st = as_stream(stream);
  const unsigned grid = (unsigned)((n + kNnT - 1) / kNnT);
  const int ki = (int)k, qi = (int)q;
  if (q <= 8) gemm_nn_kernel<8><<<grid, kNnT, 0, st>>>(A, lda, C, ldc, Y, ldy, n, ki, qi, alpha, beta);
  else if (q <= 16) gemm_nn_kernel<16><<<grid, kNnT, 0, st>>>(A, lda, C, ldc, Y, ldy, n, ki, qi, alpha, beta);
  else if (q <= 32) gemm_nn_kernel<32><<<grid, kNnT, 0, st>>>(A, lda, C, ldc, Y, ldy, n, ki, qi, alpha, beta);
  else gemm_nn_kernel<64><<<grid, kNnT, 0, st>>>(A, lda, C, ldc, Y, ldy, n, ki, qi, alpha, beta);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_chol_upper_inv_f64(const double* A, int64_t b, double* R, double* Rinv, void* stream) {
  FEM_REQUIRE(b >= 1 && b <= 64, FEM_E_INVALID, "b must be in [1, 64] (got %lld)", (long long)b);
  FEM_REQUIRE(A && R && Rinv, FEM_E_INVALID, "null pointer");
  chol_upper_inv_kernel<<<1, kCholT, 0, as_stream(stream)>>>(A, (int)b, R, Rinv);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}
