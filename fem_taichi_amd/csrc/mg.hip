// Geometric multigrid on structured node grids: the FP32 V-cycle that
// preconditions the FP64 flexible CG of the static solve (the GPU solver
// that replaces the Pardiso call of SOLVE_static.py:31-38).
//
// * Grid transfers between a node grid (x fastest) and its coarsening, which
//   keeps every second node along each axis plus the last one.  Prolongation
//   is trilinear in the index coordinate, restriction its transpose.  Both
//   are gathers (thread per output node), so no atomics and a fixed
//   summation order.
// * The damped block-Jacobi smoother x <- x + w Dinv (b - K x) and the
//   residual r = b - K x on the FP32 matrix in BSR or the reference's CSR
//   value order: 16 lanes per block row (lane b takes blocks b, b + 16, so
//   the row's contiguous run of values is read coalesced), the three row
//   sums reduced with lane shuffles.
// * y = A x for the small dense inverse of the coarsest level.
#include "common.hpp"

namespace fem {
namespace {

constexpr int kLanes = 16;      // lanes per block row (FP32 rows: 16 beats 32 and 8, measured)
constexpr int kThreads = 256;
constexpr int kRowsPerBlock = kThreads / kLanes;

// (K x)_i for block row i over the row's blocks; lanes b, b + 16, ...  Entry
// (p, q) of the row's block b sits at 9 (sm + b) + 3p + q (BSR) or at
// 9 sm + 3p co + 3b + q (the reference's CSR order).
template <int LAYOUT>
__device__ __forceinline__ void row_mul(const fem_mesh& m, const float* __restrict__ vals,
                                        const float* __restrict__ x, int64_t i, int lane, float& a0,
                                        float& a1, float& a2) {
  a0 = a1 = a2 = 0.0f;
  const int64_t sm = m.sm_node[i];
  const int co = m.co_node[i];
  const int64_t ps = LAYOUT == FEM_LAYOUT_CSR ? 3 * (int64_t)co : 3;    // stride between block rows p
#pragma unroll 2
  for (int b = lane; b < co; b += kLanes) {
    const float* __restrict__ v = LAYOUT == FEM_LAYOUT_CSR ? vals + 9 * sm + 3 * b : vals + 9 * (sm + b);
    const int64_t c = 3 * (int64_t)m.re_node[sm + b];
    const float x0 = x[c], x1 = x[c + 1], x2 = x[c + 2];
    a0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
    a1 += v[ps] * x0 + v[ps + 1] * x1 + v[ps + 2] * x2;
    a2 += v[2 * ps] * x0 + v[2 * ps + 1] * x1 + v[2 * ps + 2] * x2;
  }
#pragma unroll
  for (int o = kLanes / 2; o > 0; o >>= 1) {
    a0 += __shfl_xor(a0, o, 64);
    a1 += __shfl_xor(a1, o, 64);
    a2 += __shfl_xor(a2, o, 64);
  }
}

// MODE 0: out = b - K x.  MODE 1: out = x + w Dinv (b - K x) (out != x).
template <int MODE, int LAYOUT>
__global__ void __launch_bounds__(kThreads)
smooth_kernel(fem_mesh m, const float* __restrict__ vals, const float* __restrict__ dinv, float w,
              const float* __restrict__ x, const float* __restrict__ b, float* __restrict__ out) {
  const int lane = threadIdx.x & (kLanes - 1);
  const int64_t i = (int64_t)blockIdx.x * kRowsPerBlock + threadIdx.x / kLanes;
  // the lanes of a row take the same branch, so the shuffles see whole groups
  if (i < m.num_node) {
    float a0, a1, a2;
    row_mul<LAYOUT>(m, vals, x, i, lane, a0, a1, a2);
    if (lane == 0) {
      const float r0 = b[3 * i] - a0, r1 = b[3 * i + 1] - a1, r2 = b[3 * i + 2] - a2;
      if (MODE == 0) {
        out[3 * i] = r0;
        out[3 * i + 1] = r1;
        out[3 * i + 2] = r2;
      } else {
        const float* __restrict__ D = dinv + 9 * i;
#pragma unroll
        for (int p = 0; p < 3; ++p)
          out[3 * i + p] = x[3 * i + p] + w * (D[3 * p] * r0 + D[3 * p + 1] * r1 + D[3 * p + 2] * r2);
      }
    }
  }
}

// out = w Dinv b (the smoother's first step from x = 0).
__global__ void __launch_bounds__(kThreads)
jacobi_start_kernel(int64_t N, const float* __restrict__ dinv, float w, const float* __restrict__ b,
                    float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= N) return;
  const float* __restrict__ D = dinv + 9 * i;
  const float b0 = b[3 * i], b1 = b[3 * i + 1], b2 = b[3 * i + 2];
#pragma unroll
  for (int p = 0; p < 3; ++p) out[3 * i + p] = w * (D[3 * p] * b0 + D[3 * p + 1] * b1 + D[3 * p + 2] * b2);
}

// Weight of coarse index c for fine index f along one axis.
__device__ __forceinline__ float axis_weight(const int32_t* __restrict__ left, const float* __restrict__ wr,
                                             int f, int c) {
  const int l = left[f];
  return l == c ? 1.0f - wr[f] : (l + 1 == c ? wr[f] : 0.0f);
}

// coarse[I] = sum over fine nodes f of P[f, I] fine[f] (thread per coarse node).
// Launched on a (ceil(nc_x nb / 256), nc_y, nc_z) grid: the node's y and z
// come from the block indices, x and the column from one 32-bit division.
__global__ void __launch_bounds__(kThreads)
restrict_kernel(fem_grid_transfer t, const float* __restrict__ fine, float* __restrict__ coarse, int nb) {
  const uint32_t tx = blockIdx.x * kThreads + threadIdx.x;
  if (tx >= (uint32_t)t.nc[0] * (uint32_t)nb) return;
  const int cx = (int)(tx / (uint32_t)nb), col = (int)(tx - (uint32_t)cx * nb);
  const int cy = blockIdx.y, cz = blockIdx.z;
  const int64_t I = cx + (int64_t)t.nc[0] * (cy + (int64_t)t.nc[1] * cz);
  const int32_t* lx = t.left;
  const int32_t* ly = lx + t.nf[0];
  const int32_t* lz = ly + t.nf[1];
  const float* wx = t.wright;
  const float* wy = wx + t.nf[0];
  const float* wz = wy + t.nf[1];
  const int32_t* sx = t.span;
  const int32_t* sy = sx + 2 * t.nc[0];
  const int32_t* sz = sy + 2 * t.nc[1];
  const int64_t fx0 = t.nf[0], fxy = (int64_t)t.nf[0] * t.nf[1];
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
  const int z0 = sz[2 * cz], z1 = sz[2 * cz + 1], y0 = sy[2 * cy], y1 = sy[2 * cy + 1];
  const int x0 = sx[2 * cx], x1 = sx[2 * cx + 1];
  if (z1 - z0 <= 2 && y1 - y0 <= 2 && x1 - x0 <= 2) {
    // spans of at most three nodes (every coarsening by axis_select): all
    // 27 taps unrolled, those past a span's end clamped to it with weight 0,
    // so the loads issue together; the sums keep the loop order below
    float gx[3], gy[3], gz[3];
    int ix[3], iy[3], iz[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      iz[d] = min(z0 + d, z1);
      iy[d] = min(y0 + d, y1);
      ix[d] = min(x0 + d, x1);
      gz[d] = z0 + d <= z1 ? axis_weight(lz, wz, iz[d], cz) : 0.0f;
      gy[d] = y0 + d <= y1 ? axis_weight(ly, wy, iy[d], cy) : 0.0f;
      gx[d] = x0 + d <= x1 ? axis_weight(lx, wx, ix[d], cx) : 0.0f;
    }
#pragma unroll
    for (int dz = 0; dz < 3; ++dz)
#pragma unroll
      for (int dy = 0; dy < 3; ++dy) {
        const float gyz = gz[dz] * gy[dy];
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
          const float g = gyz * gx[dx];
          const int64_t f = 3 * (ix[dx] + fx0 * iy[dy] + fxy * iz[dz]) * nb + col;
          a0 += g * fine[f];
          a1 += g * fine[f + nb];
          a2 += g * fine[f + 2 * nb];
        }
      }
  } else {
    for (int fz = z0; fz <= z1; ++fz) {
      const float gz = axis_weight(lz, wz, fz, cz);
      for (int fy = y0; fy <= y1; ++fy) {
        const float gyz = gz * axis_weight(ly, wy, fy, cy);
        for (int fx = x0; fx <= x1; ++fx) {
          const float g = gyz * axis_weight(lx, wx, fx, cx);
          const int64_t f = 3 * (fx + fx0 * fy + fxy * fz) * nb + col;
          a0 += g * fine[f];
          a1 += g * fine[f + nb];
          a2 += g * fine[f + 2 * nb];
        }
      }
    }
  }
  const int64_t o = 3 * I * nb + col;
  coarse[o] = a0;
  coarse[o + nb] = a1;
  coarse[o + 2 * nb] = a2;
}

// fine[f] += sum over the (at most 8) coarse nodes I of P[f, I] coarse[I].
__global__ void __launch_bounds__(kThreads)
prolong_add_kernel(fem_grid_transfer t, const float* __restrict__ coarse, float* __restrict__ fine, int nb) {
  const uint32_t tx = blockIdx.x * kThreads + threadIdx.x;     // grid (ceil(nf_x nb / 256), nf_y, nf_z)
  if (tx >= (uint32_t)t.nf[0] * (uint32_t)nb) return;
  const int fx = (int)(tx / (uint32_t)nb), col = (int)(tx - (uint32_t)fx * nb);
  const int fy = blockIdx.y, fz = blockIdx.z;
  const int64_t f = fx + (int64_t)t.nf[0] * (fy + (int64_t)t.nf[1] * fz);
  const int32_t* lx = t.left;
  const int32_t* ly = lx + t.nf[0];
  const int32_t* lz = ly + t.nf[1];
  const float* wx = t.wright;
  const float* wy = wx + t.nf[0];
  const float* wz = wy + t.nf[1];
  const int cx[2] = {lx[fx], lx[fx] + 1}, cy[2] = {ly[fy], ly[fy] + 1}, cz[2] = {lz[fz], lz[fz] + 1};
  const float gx[2] = {1.0f - wx[fx], wx[fx]}, gy[2] = {1.0f - wy[fy], wy[fy]}, gz[2] = {1.0f - wz[fz], wz[fz]};
  const int64_t ncx = t.nc[0], ncxy = (int64_t)t.nc[0] * t.nc[1];
  // the right neighbour of a coinciding node has weight 0 and may lie past
  // the grid: its index is clamped and it adds zero, so all eight taps load
  // together without branches
  const int cxc[2] = {cx[0], min(cx[1], t.nc[0] - 1)}, cyc[2] = {cy[0], min(cy[1], t.nc[1] - 1)};
  const int czc[2] = {cz[0], min(cz[1], t.nc[2] - 1)};
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
#pragma unroll
  for (int kz = 0; kz < 2; ++kz)
#pragma unroll
    for (int ky = 0; ky < 2; ++ky)
#pragma unroll
      for (int kx = 0; kx < 2; ++kx) {
        const float g = gx[kx] * gy[ky] * gz[kz];
        const int64_t I = 3 * (cxc[kx] + ncx * cyc[ky] + ncxy * czc[kz]) * nb + col;
        a0 += g * coarse[I];
        a1 += g * coarse[I + nb];
        a2 += g * coarse[I + 2 * nb];
      }
  const int64_t o = 3 * f * nb + col;
  fine[o] += a0;
  fine[o + nb] += a1;
  fine[o + 2 * nb] += a2;
}

// The two transfers for a column count that is a multiple of 4: each thread
// takes 4 adjacent columns as one 16-byte load per tap and component (a
// quarter of the memory instructions of the kernels above), with every
// column's sum in the same tap order as there, so the results are the same
// bit for bit.
__device__ __forceinline__ void f4_fma(float4& a, float g, const float4 v) {
  a.x += g * v.x;
  a.y += g * v.y;
  a.z += g * v.z;
  a.w += g * v.w;
}

__global__ void __launch_bounds__(kThreads)
restrict4_kernel(fem_grid_transfer t, const float* __restrict__ fine, float* __restrict__ coarse, int nb) {
  const int nq = nb >> 2;
  const uint32_t tx = blockIdx.x * kThreads + threadIdx.x;   // grid (ceil(nc_x nb / 4 / 256), nc_y, nc_z)
  if (tx >= (uint32_t)t.nc[0] * (uint32_t)nq) return;
  const int cx = (int)(tx / (uint32_t)nq), col = 4 * (int)(tx - (uint32_t)cx * nq);
  const int cy = blockIdx.y, cz = blockIdx.z;
  const int64_t I = cx + (int64_t)t.nc[0] * (cy + (int64_t)t.nc[1] * cz);
  const int32_t* lx = t.left;
  const int32_t* ly = lx + t.nf[0];
  const int32_t* lz = ly + t.nf[1];
  const float* wx = t.wright;
  const float* wy = wx + t.nf[0];
  const float* wz = wy + t.nf[1];
  const int32_t* sx = t.span;
  const int32_t* sy = sx + 2 * t.nc[0];
  const int32_t* sz = sy + 2 * t.nc[1];
  const int64_t fx0 = t.nf[0], fxy = (int64_t)t.nf[0] * t.nf[1];
  float4 a0 = make_float4(0.f, 0.f, 0.f, 0.f), a1 = a0, a2 = a0;
  const int z0 = sz[2 * cz], z1 = sz[2 * cz + 1], y0 = sy[2 * cy], y1 = sy[2 * cy + 1];
  const int x0 = sx[2 * cx], x1 = sx[2 * cx + 1];
  auto tap = [&](int fx, int fy, int fz, float g) {
    const float* f = fine + 3 * (fx + fx0 * fy + fxy * fz) * nb + col;
    f4_fma(a0, g, *reinterpret_cast<const float4*>(f));
    f4_fma(a1, g, *reinterpret_cast<const float4*>(f + nb));
    f4_fma(a2, g, *reinterpret_cast<const float4*>(f + 2 * nb));
  };
  if (z1 - z0 <= 2 && y1 - y0 <= 2 && x1 - x0 <= 2) {
    float gx[3], gy[3], gz[3];
    int ix[3], iy[3], iz[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      iz[d] = min(z0 + d, z1);
      iy[d] = min(y0 + d, y1);
      ix[d] = min(x0 + d, x1);
      gz[d] = z0 + d <= z1 ? axis_weight(lz, wz, iz[d], cz) : 0.0f;
      gy[d] = y0 + d <= y1 ? axis_weight(ly, wy, iy[d], cy) : 0.0f;
      gx[d] = x0 + d <= x1 ? axis_weight(lx, wx, ix[d], cx) : 0.0f;
    }
#pragma unroll
    for (int dz = 0; dz < 3; ++dz)
#pragma unroll
      for (int dy = 0; dy < 3; ++dy) {
        const float gyz = gz[dz] * gy[dy];
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) tap(ix[dx], iy[dy], iz[dz], gyz * gx[dx]);
      }
  } else {
    for (int fz = z0; fz <= z1; ++fz) {
      const float gz = axis_weight(lz, wz, fz, cz);
      for (int fy = y0; fy <= y1; ++fy) {
        const float gyz = gz * axis_weight(ly, wy, fy, cy);
        for (int fx = x0; fx <= x1; ++fx) tap(fx, fy, fz, gyz * axis_weight(lx, wx, fx, cx));
      }
    }
  }
  float* o = coarse + 3 * I * nb + col;
  *reinterpret_cast<float4*>(o) = a0;
  *reinterpret_cast<float4*>(o + nb) = a1;
  *reinterpret_cast<float4*>(o + 2 * nb) = a2;
}

__global__ void __launch_bounds__(kThreads)
prolong_add4_kernel(fem_grid_transfer t, const float* __restrict__ coarse, float* __restrict__ fine, int nb) {
  const int nq = nb >> 2;
  const uint32_t tx = blockIdx.x * kThreads + threadIdx.x;   // grid (ceil(nf_x nb / 4 / 256), nf_y, nf_z)
  if (tx >= (uint32_t)t.nf[0] * (uint32_t)nq) return;
  const int fx = (int)(tx / (uint32_t)nq), col = 4 * (int)(tx - (uint32_t)fx * nq);
  const int fy = blockIdx.y, fz = blockIdx.z;
  const int64_t f = fx + (int64_t)t.nf[0] * (fy + (int64_t)t.nf[1] * fz);
  const int32_t* lx = t.left;
  const int32_t* ly = lx + t.nf[0];
  const int32_t* lz = ly + t.nf[1];
  const float* wx = t.wright;
  const float* wy = wx + t.nf[0];
  const float* wz = wy + t.nf[1];
  const int cx[2] = {lx[fx], lx[fx] + 1}, cy[2] = {ly[fy], ly[fy] + 1}, cz[2] = {lz[fz], lz[fz] + 1};
  const float gx[2] = {1.0f - wx[fx], wx[fx]}, gy[2] = {1.0f - wy[fy], wy[fy]}, gz[2] = {1.0f - wz[fz], wz[fz]};
  const int64_t ncx = t.nc[0], ncxy = (int64_t)t.nc[0] * t.nc[1];
  const int cxc[2] = {cx[0], min(cx[1], t.nc[0] - 1)}, cyc[2] = {cy[0], min(cy[1], t.nc[1] - 1)};
  const int czc[2] = {cz[0], min(cz[1], t.nc[2] - 1)};
  float4 a0 = make_float4(0.f, 0.f, 0.f, 0.f), a1 = a0, a2 = a0;
#pragma unroll
  for (int kz = 0; kz < 2; ++kz)
#pragma unroll
    for (int ky = 0; ky < 2; ++ky)
#pragma unroll
      for (int kx = 0; kx < 2; ++kx) {
        const float g = gx[kx] * gy[ky] * gz[kz];
        const float* c = coarse + 3 * (cxc[kx] + ncx * cyc[ky] + ncxy * czc[kz]) * nb + col;
        f4_fma(a0, g, *reinterpret_cast<const float4*>(c));
        f4_fma(a1, g, *reinterpret_cast<const float4*>(c + nb));
        f4_fma(a2, g, *reinterpret_cast<const float4*>(c + 2 * nb));
      }
  float* o = fine + 3 * f * nb + col;
  float4 v0 = *reinterpret_cast<float4*>(o), v1 = *reinterpret_cast<float4*>(o + nb);
  float4 v2 = *reinterpret_cast<float4*>(o + 2 * nb);
  v0.x += a0.x; v0.y += a0.y; v0.z += a0.z; v0.w += a0.w;
  v1.x += a1.x; v1.y += a1.y; v1.z += a1.z; v1.w += a1.w;
  v2.x += a2.x; v2.y += a2.y; v2.z += a2.z; v2.w += a2.w;
  *reinterpret_cast<float4*>(o) = v0;
  *reinterpret_cast<float4*>(o + nb) = v1;
  *reinterpret_cast<float4*>(o + 2 * nb) = v2;
}

// ---- several right-hand sides at once (the inner solves of block Lanczos):
// blocks of vectors are node-interleaved with columns fastest,
// X[(3 i + p) nb + c], as in the multi-RHS PCG.  Thread per (node, column).

// MODE 0: OUT = B - K X.  MODE 1: OUT = X + w Dinv (B - K X).  MODE 2: OUT = w Dinv B.
template <int MODE>
__global__ void __launch_bounds__(kThreads)
smooth_multi_kernel(fem_mesh m, const float* __restrict__ vals, const float* __restrict__ dinv, float w,
                    const float* __restrict__ X, const float* __restrict__ B, float* __restrict__ OUT, int nb) {
  const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t >= m.num_node * nb) return;
  // (node, column) of the thread; a 32-bit division whenever the count fits
  const int64_t i = t < ((int64_t)1 << 32) ? (int64_t)((uint32_t)t / (uint32_t)nb) : t / nb;
  const int c = (int)(t - i * nb);
  const int64_t o = 3 * i * nb + c;
  float r0 = B[o], r1 = B[o + nb], r2 = B[o + 2 * nb];
  if (MODE != 2) {
    const int64_t sm = m.sm_node[i];
    const int co = m.co_node[i];
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll 4   // more gathers in flight (1.55 -> 1.31 ms at 8 columns on 2.3M nodes)
    for (int b = 0; b < co; ++b) {
      const float* __restrict__ v = vals + 9 * (sm + b);
      const int64_t xc = 3 * (int64_t)m.re_node[sm + b] * nb + c;
      const float x0 = X[xc], x1 = X[xc + nb], x2 = X[xc + 2 * nb];
      a0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
      a1 += v[3] * x0 + v[4] * x1 + v[5] * x2;
      a2 += v[6] * x0 + v[7] * x1 + v[8] * x2;
    }
    r0 -= a0;
    r1 -= a1;
    r2 -= a2;
  }
  if (MODE == 0) {
    OUT[o] = r0;
    OUT[o + nb] = r1;
    OUT[o + 2 * nb] = r2;
    return;
  }
  const float* __restrict__ D = dinv + 9 * i;
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const float d = w * (D[3 * p] * r0 + D[3 * p + 1] * r1 + D[3 * p + 2] * r2);
    OUT[o + p * nb] = MODE == 1 ? X[o + p * nb] + d : d;
  }
}

// The same passes for 16 columns (MODE 0 and 1): the 8 lanes of a row take
// two columns each and stage the row's K blocks in LDS eight at a time with
// coalesced loads, instead of every lane loading every block (16 lanes per
// row read the same 36 bytes).  Same sums in the same order per column as
// smooth_multi_kernel (bit-identical); on the 2.3M-node blade's finest level
// 1.63 ms instead of 1.97 ms (at 8 columns no gain).
constexpr int kSm16Team = 8, kSm16Rows = 64 / kSm16Team, kSm16Stride = 9 * kSm16Team + 1;
template <int MODE>
__global__ void __launch_bounds__(kThreads)
smooth_multi16_kernel(fem_mesh m, const float* __restrict__ vals, const float* __restrict__ dinv, float w,
                      const float* __restrict__ X, const float* __restrict__ B, float* __restrict__ OUT) {
  constexpr int nb = 16;
  __shared__ float sk[kThreads / 64][kSm16Rows][kSm16Stride];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int r = l / kSm16Team, j = l % kSm16Team;
  const int64_t i = ((int64_t)blockIdx.x * kThreads + threadIdx.x) / kSm16Team;
  if (i >= m.num_node) return;       // whole row teams leave together
  const int64_t sm = m.sm_node[i];
  const int co = m.co_node[i];
  const int c = 2 * j;
  float a0 = 0.f, a1 = 0.f, a2 = 0.f, b0 = 0.f, b1 = 0.f, b2 = 0.f;
  float* s = &sk[wv][r][0];
  for (int c0 = 0; c0 < co; c0 += kSm16Team) {
    const int nblk = min(kSm16Team, co - c0);
    const float* src = vals + 9 * (sm + c0);
    float kv[9];                    // the nine loads in flight together, then the LDS stores
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int d = j + kSm16Team * k;
      kv[k] = d < 9 * nblk ? src[d] : float(0);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int d = j + kSm16Team * k;
      if (d < 9 * nblk) s[d] = kv[k];
    }
    const int colj = j < nblk ? m.re_node[sm + c0 + j] : 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll 4
    for (int b = 0; b < kSm16Team; ++b) {
      if (b < nblk) {
        const float* v = s + 9 * b;
        const int64_t xc = 3 * (int64_t)__shfl(colj, b, kSm16Team) * nb + c;
        const float2 x0 = *reinterpret_cast<const float2*>(X + xc);
        const float2 x1 = *reinterpret_cast<const float2*>(X + xc + nb);
        const float2 x2 = *reinterpret_cast<const float2*>(X + xc + 2 * nb);
        a0 += v[0] * x0.x + v[1] * x1.x + v[2] * x2.x;
        a1 += v[3] * x0.x + v[4] * x1.x + v[5] * x2.x;
        a2 += v[6] * x0.x + v[7] * x1.x + v[8] * x2.x;
        b0 += v[0] * x0.y + v[1] * x1.y + v[2] * x2.y;
        b1 += v[3] * x0.y + v[4] * x1.y + v[5] * x2.y;
        b2 += v[6] * x0.y + v[7] * x1.y + v[8] * x2.y;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  const float acc[2][3] = {{a0, a1, a2}, {b0, b1, b2}};
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t o = 3 * i * nb + c + h;
    const float r0 = B[o] - acc[h][0], r1 = B[o + nb] - acc[h][1], r2 = B[o + 2 * nb] - acc[h][2];
    if (MODE == 0) {
      OUT[o] = r0;
      OUT[o + nb] = r1;
      OUT[o + 2 * nb] = r2;
    } else {
      const float* __restrict__ D = dinv + 9 * i;
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        const float d = w * (D[3 * p] * r0 + D[3 * p + 1] * r1 + D[3 * p + 2] * r2);
        OUT[o + p * nb] = X[o + p * nb] + d;
      }
    }
  }
}

// Y = A X, A dense n x n row-major (the coarsest level's inverse), X and Y
// n x nb with the columns fastest.  A work-group computes a 4-row x
// 16-column tile of Y over k in chunks of 256: each chunk's A rows (4 x 256,
// coalesced) and X rows (256 x 16) go to LDS, and the next chunk is already
// loaded into registers while this one is summed.  The 256 threads are 4
// slices (one wave each) of 64 (row, column) pairs; slice s sums the k in
// [64 s, 64 s + 64) of each chunk in ascending order, and the four slices'
// sums are added in slice order at the end (deterministic).  ceil(n / 4)
// work-groups spread the coarsest levels' few thousand rows over the CUs
// (2,970 rows and 16 columns: 37 us, against 71 us with 16-row tiles, a
// quarter as many work-groups, each with the same chain of n / 256 chunks).
constexpr int kGemmRows = 4, kGemmCols = 16, kGemmK = 256, kGemmSlices = 4;
constexpr int kGemmA = kGemmRows * kGemmK / kThreads, kGemmX = kGemmK * kGemmCols / kThreads;   // 4, 16
__global__ void __launch_bounds__(kThreads)
gemm_nb_kernel(const float* __restrict__ A, int64_t n, const float* __restrict__ X, float* __restrict__ Y, int nb) {
  __shared__ float sA[kGemmRows * (kGemmK + 1)];
  __shared__ float sX[kGemmK * kGemmCols];
  __shared__ float sR[kGemmSlices * kGemmRows * kGemmCols];
  const int tid = threadIdx.x;
  const int sl = tid / (kGemmRows * kGemmCols), rc = tid % (kGemmRows * kGemmCols);
  const int r = rc / kGemmCols, c = rc % kGemmCols;
  const int64_t row0 = (int64_t)blockIdx.x * kGemmRows;
  const int col0 = blockIdx.y * kGemmCols;
  constexpr int kSlice = kGemmK / kGemmSlices;
  float ra[kGemmA], rx[kGemmX];
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int i = 0; i < kGemmA; ++i) {
      const int e = tid + i * kThreads, ar = e / kGemmK, ak = e % kGemmK;
      const int64_t gr = row0 + ar, gk = k0 + ak;
      ra[i] = (gr < n && gk < n) ? A[gr * n + gk] : 0.f;
    }
#pragma unroll
    for (int i = 0; i < kGemmX; ++i) {
      const int e = tid + i * kThreads, xk = e / kGemmCols, xc = e % kGemmCols;
      const int64_t gk = k0 + xk;
      rx[i] = (gk < n && col0 + xc < nb) ? X[gk * nb + col0 + xc] : 0.f;
    }
  };
  float s = 0.f;
  load(0);
  for (int64_t k0 = 0; k0 < n; k0 += kGemmK) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kGemmA; ++i) {
      const int e = tid + i * kThreads;
      sA[(e / kGemmK) * (kGemmK + 1) + e % kGemmK] = ra[i];
    }
#pragma unroll
    for (int i = 0; i < kGemmX; ++i) sX[tid + i * kThreads] = rx[i];
    __syncthreads();
    if (k0 + kGemmK < n) load(k0 + kGemmK);
    const float* a = sA + r * (kGemmK + 1) + sl * kSlice;
    const float* x = sX + sl * kSlice * kGemmCols + c;
#pragma unroll 16
    for (int k = 0; k < kSlice; ++k) s = fmaf(a[k], x[k * kGemmCols], s);
  }
  sR[tid] = s;
  __syncthreads();
  if (sl == 0) {
    float t = sR[rc];
#pragma unroll
    for (int q = 1; q < kGemmSlices; ++q) t += sR[q * kGemmRows * kGemmCols + rc];
    const int64_t row = row0 + r;
    if (row < n && col0 + c < nb) Y[row * nb + col0 + c] = t;
  }
}

// y = A x, A dense n x n row-major; one wave per row.
__global__ void __launch_bounds__(kThreads)
gemv_kernel(const float* __restrict__ A, int64_t n, const float* __restrict__ x, float* __restrict__ y) {
  const int64_t row = (int64_t)blockIdx.x * (kThreads / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  float s = 0.0f;
  if (row < n) {
    const float* __restrict__ a = A + row * n;
    for (int64_t c = lane; c < n; c += 64) s += a[c] * x[c];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (row < n && lane == 0) y[row] = s;
}

int check_transfer(const fem_grid_transfer* t) {
  FEM_REQUIRE(t && t->left && t->wright && t->span, FEM_E_INVALID, "null grid transfer table");
  for (int a = 0; a < 3; ++a) {
    FEM_REQUIRE(t->nf[a] >= 1 && t->nc[a] >= 1 && t->nc[a] <= t->nf[a], FEM_E_INVALID,
                "grid transfer axis %d: %d fine, %d coarse nodes", a, t->nf[a], t->nc[a]);
  }
  return FEM_OK;
}

}  // namespace
}  // namespace fem

using namespace fem;

extern "C" int fem_mg_smooth(const fem_mesh* mesh, const float* vals, int layout, const float* dinv, float omega,
                             const float* x, const float* b, float* out, int mode, void* stream) {
  FEM_REQUIRE(mesh && b && out, FEM_E_INVALID, "null pointer");
  FEM_REQUIRE(mode >= 0 && mode <= 2, FEM_E_INVALID, "mode must be 0 (residual), 1 (update) or 2 (start)");
  FEM_REQUIRE(layout == FEM_LAYOUT_BSR || layout == FEM_LAYOUT_CSR, FEM_E_INVALID, "bad layout %d", layout);
  FEM_REQUIRE(mode == 0 || dinv, FEM_E_INVALID, "null dinv");
  FEM_REQUIRE(mode == 2 || (vals && x), FEM_E_INVALID, "null vals or x");
  FEM_REQUIRE(mode != 1 || out != x, FEM_E_INVALID, "the update needs out != x (x is gathered)");
  const int64_t N = mesh->num_node;
  if (N == 0) return FEM_OK;
  hipStream_t st = as_stream(stream);
  if (mode == 2) {
    jacobi_start_kernel<<<ceil_div(N, kThreads), kThreads, 0, st>>>(N, dinv, omega, b, out);
  } else {
    const int64_t nb = ceil_div(N, kRowsPerBlock);
    if (mode == 0 && layout == FEM_LAYOUT_BSR)
      smooth_kernel<0, FEM_LAYOUT_BSR><<<nb, kThreads, 0, st>>>(*mesh, vals, dinv, omega, x, b, out);
    else if (mode == 0)
      smooth_kernel<0, FEM_LAYOUT_CSR><<<nb, kThreads, 0, st>>>(*mesh, vals, dinv, omega, x, b, out);
    else if (layout == FEM_LAYOUT_BSR)
      smooth_kernel<1, FEM_LAYOUT_BSR><<<nb, kThreads, 0, st>>>(*mesh, vals, dinv, omega, x, b, out);
    else
      smooth_kernel<1, FEM_LAYOUT_CSR><<<nb, kThreads, 0, st>>>(*mesh, vals, dinv, omega, x, b, out);
  }
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_mg_restrict(const fem_grid_transfer* t, const float* fine, float* coarse, int64_t nrhs,
                               void* stream) {
  int rc = check_transfer(t);
  if (rc) return rc;
  FEM_REQUIRE(fine && coarse, FEM_E_INVALID, "null vector");
  FEM_REQUIRE(nrhs >= 1 && nrhs <= 64, FEM_E_INVALID, "nrhs must be in [1, 64]");
  const int64_t Nc = (int64_t)t->nc[0] * t->nc[1] * t->nc[2];
  if (Nc == 0) return FEM_OK;
  FEM_REQUIRE(t->nc[1] <= 65535 && t->nc[2] <= 65535, FEM_E_LIMIT, "coarse grid beyond 65535 nodes along y or z");
  if (nrhs % 4 == 0 && reinterpret_cast<uintptr_t>(fine) % 16 == 0 && reinterpret_cast<uintptr_t>(coarse) % 16 == 0) {
    const dim3 grid((unsigned)ceil_div((int64_t)t->nc[0] * (nrhs / 4), kThreads), (unsigned)t->nc[1],
                    (unsigned)t->nc[2]);
    restrict4_kernel<<<grid, kThreads, 0, as_stream(stream)>>>(*t, fine, coarse, (int)nrhs);
    FEM_LAUNCH_CHECK();
    return FEM_OK;
  }
  const dim3 grid((unsigned)ceil_div((int64_t)t->nc[0] * nrhs, kThreads), (unsigned)t->nc[1], (unsigned)t->nc[2]);
  restrict_kernel<<<grid, kThreads, 0, as_stream(stream)>>>(*t, fine, coarse, (int)nrhs);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_mg_prolong_add(const fem_grid_transfer* t, const float* coarse, float* fine, int64_t nrhs,
                                  void* stream) {
  int rc = check_transfer(t);
  if (rc) return rc;
  FEM_REQUIRE(fine && coarse, FEM_E_INVALID, "null vector");
  FEM_REQUIRE(nrhs >= 1 && nrhs <= 64, FEM_E_INVALID, "nrhs must be in [1, 64]");
  const int64_t Nf = (int64_t)t->nf[0] * t->nf[1] * t->nf[2];
  if (Nf == 0) return FEM_OK;
  FEM_REQUIRE(t->nf[1] <= 65535 && t->nf[2] <= 65535, FEM_E_LIMIT, "fine grid beyond 65535 nodes along y or z");
  if (nrhs % 4 == 0 && reinterpret_cast<uintptr_t>(fine) % 16 == 0 && reinterpret_cast<uintptr_t>(coarse) % 16 == 0) {
    const dim3 grid((unsigned)ceil_div((int64_t)t->nf[0] * (nrhs / 4), kThreads), (unsigned)t->nf[1],
                    (unsigned)t->nf[2]);
    prolong_add4_kernel<<<grid, kThreads, 0, as_stream(stream)>>>(*t, coarse, fine, (int)nrhs);
    FEM_LAUNCH_CHECK();
    return FEM_OK;
  }
  const dim3 grid((unsigned)ceil_div((int64_t)t->nf[0] * nrhs, kThreads), (unsigned)t->nf[1], (unsigned)t->nf[2]);
  prolong_add_kernel<<<grid, kThreads, 0, as_stream(stream)>>>(*t, coarse, fine, (int)nrhs);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_mg_smooth_multi(const fem_mesh* mesh, const float* vals, const float* dinv, float omega,
                                   const float* X, const float* B, float* OUT, int64_t nrhs, int mode,
                                   void* stream) {
  FEM_REQUIRE(mesh && B && OUT, FEM_E_INVALID, "null pointer");
  FEM_REQUIRE(nrhs >= 1 && nrhs <= 64, FEM_E_INVALID, "nrhs must be in [1, 64]");
  FEM_REQUIRE(mode >= 0 && mode <= 2, FEM_E_INVALID, "mode must be 0 (residual), 1 (update) or 2 (start)");
  FEM_REQUIRE(mode == 0 || dinv, FEM_E_INVALID, "null dinv");
  FEM_REQUIRE(mode == 2 || (vals && X), FEM_E_INVALID, "null vals or X");
  FEM_REQUIRE(mode != 1 || OUT != X, FEM_E_INVALID, "the update needs OUT != X (X is gathered)");
  const int64_t T = mesh->num_node * nrhs;
  if (T == 0) return FEM_OK;
  hipStream_t st = as_stream(stream);
  const int nb = (int)nrhs;
  // 16 columns, residual or update: K staged in LDS, two columns per lane.
  // Otherwise one column per thread: measured on the 2.3M-node blade against
  // two and four columns per thread (residual mode, unrolled block loop), 8
  // columns take 1.31 ms.
  if (nb == 16 && mode != 2 && reinterpret_cast<uintptr_t>(X) % 8 == 0) {
    const int64_t T16 = mesh->num_node * kSm16Team;
    if (mode == 0)
      smooth_multi16_kernel<0><<<ceil_div(T16, kThreads), kThreads, 0, st>>>(*mesh, vals, dinv, omega, X, B, OUT);
    else
      smooth_multi16_kernel<1><<<ceil_div(T16, kThreads), kThreads, 0, st>>>(*mesh, vals, dinv, omega, X, B, OUT);
    FEM_LAUNCH_CHECK();
    return FEM_OK;
  }
  if (mode == 0)
    smooth_multi_kernel<0><<<ceil_div(T, kThreads), kThreads, 0, st>>>(*mesh, vals, dinv, omega, X, B, OUT, nb);
  else if (mode == 1)
    smooth_multi_kernel<1><<<ceil_div(T, kThreads), kThreads, 0, st>>>(*mesh, vals, dinv, omega, X, B, OUT, nb);
  else
    smooth_multi_kernel<2><<<ceil_div(T, kThreads), kThreads, 0, st>>>(*mesh, vals, dinv, omega, X, B, OUT, nb);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_dense_gemm_nb_f32(const float* A, int64_t n, const float* X, float* Y, int64_t nrhs,
                                     void* stream) {
  FEM_REQUIRE(A && X && Y && n >= 0, FEM_E_INVALID, "null pointer or negative size");
  FEM_REQUIRE(nrhs >= 1 && nrhs <= 64, FEM_E_INVALID, "nrhs must be in [1, 64]");
  FEM_REQUIRE(X != Y, FEM_E_INVALID, "Y must not alias X");
  if (n == 0) return FEM_OK;
  const dim3 grid((unsigned)ceil_div(n, kGemmRows), (unsigned)ceil_div(nrhs, kGemmCols));
  gemm_nb_kernel<<<grid, kThreads, 0, as_stream(stream)>>>(A, n, X, Y, (int)nrhs);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_dense_gemv_f32(const float* A, int64_t n, const float* x, float* y, void* stream) {
  FEM_REQUIRE(A && x && y && n >= 0, FEM_E_INVALID, "null pointer or negative size");
  FEM_REQUIRE(x != y, FEM_E_INVALID, "y must not alias x");
  if (n == 0) return FEM_OK;
  gemv_kernel<<<ceil_div(n, kThreads / 64), kThreads, 0, as_stream(stream)>>>(A, n, x, y);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}
