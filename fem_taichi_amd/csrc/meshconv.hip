// 3x3x3 mesh convolution on the structured node grid (the learned
// preconditioner's layers), BF16 operands on MFMA with FP32 accumulation.
//
// Layout: activations are channels-last, x[p * C + c], p = ix + nx*(iy +
// ny*iz) (the node index of a structured grid), C padded to a multiple of 32
// (inputs) / 16 (outputs); weights w[tap][co][ci] with ci contiguous, tap =
// (dz+1)*9 + (dy+1)*3 + (dx+1); out-of-grid neighbours read as zero.
//
// Forward (implicit GEMM on v_mfma_f32_16x16x32_bf16, LDS-tiled: see
// conv3_fwd_tiled_kernel).  Operand layout of that instruction: lane l
// supplies A[l%16][8*(l/16)+i] and B[8*(l/16)+i][l%16], i < 8, and receives
// D[4*(l/16)+i][l%16], i < 4.
// Backward data is the forward kernel on flipped, channel-transposed
// weights.  The weight gradient is a GEMM over pixels: each work-group owns
// (tap, pixel chunk), stages 32-pixel tiles of dy and of the shifted input in
// LDS transposed (so each lane's 8 pixels are contiguous), and writes a
// partial; partials are summed over chunks in index order (deterministic).
#include "common.hpp"

namespace fem {
namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kCT = 256;

__device__ __forceinline__ bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int i = 0; i < 8; ++i) z[i] = (bf16)0.0f;
  return z;
}

// ---- LDS-tiled forward: a work-group owns a 16 x 4 x 4 block of output
// nodes (x fastest); wave w owns the z-plane w of it, i.e. 4 rows of 16 nodes,
// and all cout channels: 4 x (cout / 16) accumulator tiles, so each A
// fragment feeds cout / 16 MFMAs and each B fragment 4.  The K loop runs
// chunk-outer: per 32-channel chunk the 18 x 6 x 6 halo of input nodes is
// staged in LDS once, the next chunk's halo loaded into registers while this
// one is multiplied (rows of 32 channels padded to 40, which keeps the
// 16-byte fragment reads of 16 consecutive nodes conflict-free), then the 27
// taps run two at a time: each pair's weight slices (cout x 32 channels each)
// are staged in LDS for all four waves, the next pair loaded into registers
// while the current one is multiplied, one barrier per pair.
// The epilogue stages the tile's outputs in LDS and writes whole node rows.
constexpr int kTX = 16, kTY = 4, kTZ = 4;
constexpr int kHX = kTX + 2, kHY = kTY + 2, kHZ = kTZ + 2;
constexpr int kHalo = kHX * kHY * kHZ;                  // 648 halo nodes
constexpr int kLD = 40;                                 // bf16 per staged row (32 + 8 pad)
constexpr int kTileNodes = kTX * kTY * kTZ;             // 256
constexpr int kLdsA = kHalo * kLD * 2;                  // 51,840 B
constexpr int kLdsOut = kTileNodes * 64 * 4;            // f32 staging of 64 channels
constexpr int kLdsB = 4 * 64 * kLD * 2;                 // two pairs of weight slices
constexpr int kLdsTiled = (kLdsA + kLdsB > kLdsOut) ? kLdsA + kLdsB : kLdsOut;

template <int NT>
__global__ void __launch_bounds__(kCT)
conv3_fwd_tiled_kernel(const bf16* __restrict__ x, int cin, const bf16* __restrict__ w,
                       const float* __restrict__ bias, int cout, int nx, int ny, int nz,
                       bf16* __restrict__ y_act, bf16* __restrict__ y_pre, float* __restrict__ y_f32, int relu) {
  __shared__ __align__(16) unsigned char smem[kLdsTiled];
  bf16* sA = reinterpret_cast<bf16*>(smem);
  bf16* sB = reinterpret_cast<bf16*>(smem + kLdsA);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int bx = blockIdx.x * kTX, by = blockIdx.y * kTY, bz = blockIdx.z * kTZ;
  f32x4 acc[kTY][NT];
#pragma unroll
  for (int r = 0; r < kTY; ++r)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[r][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  // weight slices (tap, c0) are staged in LDS two taps at a time, double
  // buffered: slice row co < cout of 4 pieces of 8 channels, thread -> one
  // piece; the next pair is loaded into registers while the current pair is
  // multiplied, so one barrier serves two taps.
  const int brow = tid >> 2, bpart = tid & 3;
  const bool bload = brow < cout;
  constexpr int kStage = (kHalo * 4 + kCT - 1) / kCT;    // halo pieces per thread
  const int64_t wtap = (int64_t)cout * cin;              // stride between taps
  auto slice = [&](int pb, int k) { return sB + (pb * 2 + k) * (64 * kLD); };
  // this thread's share of a chunk's halo, loaded into registers (all loads
  // in flight together); the next chunk's share is loaded during this chunk's taps
  bf16x8 hv[kStage];
  auto load_halo = [&](int c) {
#pragma unroll
    for (int q = 0; q < kStage; ++q) {
      const int e = tid + q * kCT;
      hv[q] = zero8();
      if (e < kHalo * 4) {
        const int hn = e >> 2, part = e & 3;
        const int hx = hn % kHX, hy = (hn / kHX) % kHY, hz = hn / (kHX * kHY);
        const int gx = bx + hx - 1, gy = by + hy - 1, gz = bz + hz - 1;
        if (gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz)
          hv[q] = *reinterpret_cast<const bf16x8*>(x + (((int64_t)gz * ny + gy) * nx + gx) * cin + c + part * 8);
      }
    }
  };
  load_halo(0);
  for (int c0 = 0; c0 < cin; c0 += 32) {
    const bf16* wc = w + (int64_t)brow * cin + c0 + bpart * 8;
    const bf16x8 b0 = bload ? *reinterpret_cast<const bf16x8*>(wc) : zero8();
    const bf16x8 b1 = bload ? *reinterpret_cast<const bf16x8*>(wc + wtap) : zero8();
    __syncthreads();   // the previous chunk's readers of sA / sB are done
#pragma unroll
    for (int q = 0; q < kStage; ++q) {
      const int e = tid + q * kCT;
      if (e < kHalo * 4) *reinterpret_cast<bf16x8*>(sA + (e >> 2) * kLD + (e & 3) * 8) = hv[q];
    }
    if (bload) {
      *reinterpret_cast<bf16x8*>(slice(0, 0) + brow * kLD + bpart * 8) = b0;
      *reinterpret_cast<bf16x8*>(slice(0, 1) + brow * kLD + bpart * 8) = b1;
    }
    __syncthreads();
    if (c0 + 32 < cin) load_halo(c0 + 32);
    for (int tp = 0; tp < 27; tp += 2) {
      const int pb = (tp >> 1) & 1;
      bf16x8 n0 = zero8(), n1 = zero8();
      if (bload && tp + 2 < 27) n0 = *reinterpret_cast<const bf16x8*>(wc + (tp + 2) * wtap);
      if (bload && tp + 3 < 27) n1 = *reinterpret_cast<const bf16x8*>(wc + (tp + 3) * wtap);
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int tap = tp + k;
        if (tap >= 27) break;
        const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
        const bf16* sBb = slice(pb, k);
        bf16x8 Bf[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) Bf[t] = *reinterpret_cast<const bf16x8*>(sBb + (t * 16 + m) * kLD + kq * 8);
#pragma unroll
        for (int r = 0; r < kTY; ++r) {
          const int hn = ((wave + 1 + dz) * kHY + (r + 1 + dy)) * kHX + (m + 1 + dx);
          const bf16x8 Af = *reinterpret_cast<const bf16x8*>(sA + hn * kLD + kq * 8);
#pragma unroll
          for (int t = 0; t < NT; ++t)
            acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Af, Bf[t], acc[r][t], 0, 0, 0);
        }
      }
      if (tp + 2 < 27) {
        if (bload) {
          *reinterpret_cast<bf16x8*>(slice(pb ^ 1, 0) + brow * kLD + bpart * 8) = n0;
          *reinterpret_cast<bf16x8*>(slice(pb ^ 1, 1) + brow * kLD + bpart * 8) = n1;
        }
        __syncthreads();   // the next pair visible; this pair's readers done before it is overwritten
      }
    }
  }
  // ---- epilogue: bias, stage the tile (node-major, cout contiguous) in LDS, write node rows
  __syncthreads();
  float* sOf = reinterpret_cast<float*>(smem);
  bf16* sOb = reinterpret_cast<bf16*>(smem);
  const bool f32out = y_f32 != nullptr;
#pragma unroll
  for (int r = 0; r < kTY; ++r)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int ch = t * 16 + m;
      const float b = bias ? bias[ch] : 0.f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int node = (kq * 4 + i) + kTX * (r + kTY * wave);   // tile node index, x fastest
        const float v = acc[r][t][i] + b;
        if (f32out) sOf[node * cout + ch] = v;
        else sOb[node * cout + ch] = (bf16)v;
      }
    }
  __syncthreads();
  // node (tx, ty, tz) of the tile -> global node; cout values contiguous per node
  for (int e = tid; e < kTileNodes * cout; e += kCT) {
    const int node = e / cout, ch = e - node * cout;
    const int tx = node % kTX, ty = (node / kTX) % kTY, tz = node / (kTX * kTY);
    const int gx = bx + tx, gy = by + ty, gz = bz + tz;
    if (gx >= nx || gy >= ny || gz >= nz) continue;
    const int64_t o = (((int64_t)gz * ny + gy) * nx + gx) * cout + ch;
    if (f32out) {
      const float v = sOf[e];
      y_f32[o] = v;
      if (y_pre) y_pre[o] = (bf16)v;
      if (y_act) y_act[o] = (bf16)((relu && v < 0.f) ? 0.f : v);
    } else {
      const bf16 v = sOb[e];
      if (y_pre) y_pre[o] = v;
      if (y_act) y_act[o] = (relu && (float)v < 0.f) ? (bf16)0.0f : v;
    }
  }
}

// ---- first layer with its 4 input channels packed along K: k = 4 tap + ci
// (108 of 128 used), so a 16-node tile takes 4 MFMA k-steps instead of the
// padded kernel's 27 taps of 32 channels (6.75x fewer MFMAs, and 8 bytes of
// input per node instead of 64).  x4 is (P, 4) bf16; wp[s][co][kk] (4 x cout
// x 32 bf16) holds w[tap][co][ci] at 32 s + kk = 4 tap + ci, zero beyond tap
// 26.  Same tile as the padded kernel (16 x 4 x 4 nodes, wave w = z-plane w):
// lane (m, kq) of k-step s supplies the 4 channels of taps 8 s + 2 kq and
// 8 s + 2 kq + 1 at node m, read as two 8-byte halo words from LDS; the B
// fragments of all four k-steps stay in registers.
constexpr int kLdsC4 = kTileNodes * 64 * 2;             // bf16 output staging (the halo fits below it)
static_assert(kHalo * 8 <= kLdsC4, "the packed halo must fit in the output staging area");

template <int NT>
__global__ void __launch_bounds__(kCT)
conv3_fwd_c4_kernel(const bf16* __restrict__ x4, const bf16* __restrict__ wp, const float* __restrict__ bias,
                    int cout, int nx, int ny, int nz, bf16* __restrict__ y_act, bf16* __restrict__ y_pre,
                    int relu) {
  __shared__ __align__(16) unsigned char smem[kLdsC4];
  uint2* sH = reinterpret_cast<uint2*>(smem);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int bx = blockIdx.x * kTX, by = blockIdx.y * kTY, bz = blockIdx.z * kTZ;
  for (int hn = tid; hn < kHalo; hn += kCT) {
    const int hx = hn % kHX, hy = (hn / kHX) % kHY, hz = hn / (kHX * kHY);
    const int gx = bx + hx - 1, gy = by + hy - 1, gz = bz + hz - 1;
    uint2 v = make_uint2(0u, 0u);
    if (gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz)
      v = *reinterpret_cast<const uint2*>(x4 + (((int64_t)gz * ny + gy) * nx + gx) * 4);
    sH[hn] = v;
  }
  bf16x8 Bf[4][NT];
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int t = 0; t < NT; ++t)
      Bf[s][t] = *reinterpret_cast<const bf16x8*>(wp + ((int64_t)s * cout + t * 16 + m) * 32 + kq * 8);
  f32x4 acc[kTY][NT];
#pragma unroll
  for (int r = 0; r < kTY; ++r)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[r][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  __syncthreads();
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int ta = 8 * s + 2 * kq, tb = ta + 1;
    const int dza = ta / 9 - 1, dya = (ta / 3) % 3 - 1, dxa = ta % 3 - 1;
    const int dzb = tb / 9 - 1, dyb = (tb / 3) % 3 - 1, dxb = tb % 3 - 1;
#pragma unroll
    for (int r = 0; r < kTY; ++r) {
      uint2 va = make_uint2(0u, 0u), vb = make_uint2(0u, 0u);
      if (ta < 27) va = sH[((wave + 1 + dza) * kHY + (r + 1 + dya)) * kHX + (m + 1 + dxa)];
      if (tb < 27) vb = sH[((wave + 1 + dzb) * kHY + (r + 1 + dyb)) * kHX + (m + 1 + dxb)];
      const uint4 a4 = make_uint4(va.x, va.y, vb.x, vb.y);
      bf16x8 Af;
      __builtin_memcpy(&Af, &a4, sizeof(Af));
#pragma unroll
      for (int t = 0; t < NT; ++t)
        acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Af, Bf[s][t], acc[r][t], 0, 0, 0);
    }
  }
  // epilogue: bias, stage the tile (node-major, cout contiguous) in LDS, write node rows
  __syncthreads();
  bf16* sOb = reinterpret_cast<bf16*>(smem);
#pragma unroll
  for (int r = 0; r < kTY; ++r)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int ch = t * 16 + m;
      const float b = bias ? bias[ch] : 0.f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int node = (kq * 4 + i) + kTX * (r + kTY * wave);
        sOb[node * cout + ch] = (bf16)(acc[r][t][i] + b);
      }
    }
  __syncthreads();
  for (int e = tid; e < kTileNodes * cout; e += kCT) {
    const int node = e / cout, ch = e - node * cout;
    const int tx = node % kTX, ty = (node / kTX) % kTY, tz = node / (kTX * kTY);
    const int gx = bx + tx, gy = by + ty, gz = bz + tz;
    if (gx >= nx || gy >= ny || gz >= nz) continue;
    const int64_t o = (((int64_t)gz * ny + gy) * nx + gx) * cout + ch;
    const bf16 v = sOb[e];
    if (y_pre) y_pre[o] = v;
    if (y_act) y_act[o] = (relu && (float)v < 0.f) ? (bf16)0.0f : v;
  }
}

// wp[s][co][kk] = w[tap][co][ci] with 32 s + kk = 4 tap + ci, tap < 27, ci < 4 (else 0)
__global__ void pack_c4_kernel(const bf16* __restrict__ w, int cin_w, int cout, bf16* __restrict__ wp) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * cout * 32) return;
  const int s = t / (cout * 32), co = (t / 32) % cout, kk = t % 32;
  const int k = 32 * s + kk, tap = k >> 2, ci = k & 3;
  wp[t] = tap < 27 ? w[((int64_t)tap * cout + co) * cin_w + ci] : (bf16)0.0f;
}

// x4[p][c] = bf16(float(r[3p + c] / s)) for c < 3, bf16(efeat[p]) for c = 3; s = *scale
__global__ void features_c4_kernel(const double* __restrict__ r, int64_t P, const double* __restrict__ scale,
                                   const float* __restrict__ efeat, bf16* __restrict__ x4) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * P) return;
  const int64_t p = t >> 2;
  const int c = (int)(t & 3);
  x4[t] = c < 3 ? (bf16)(float)(r[3 * p + c] / scale[0]) : (bf16)efeat[p];
}

// c[3p + j] = (double)y[p * ld + j] * (s / e_ref) * mask[3p + j] (mask may be NULL); s = *scale
__global__ void correction_kernel(const float* __restrict__ y, int ld, int64_t P, const double* __restrict__ scale,
                                  double e_ref, const double* __restrict__ mask, double* __restrict__ c) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * P) return;
  const int64_t p = t / 3;
  const int j = (int)(t - 3 * p);
  const double v = (double)y[p * ld + j] * (scale[0] / e_ref);
  c[t] = mask ? v * mask[t] : v;
}

// wt[tap][ci][co] = w[26 - tap][co][ci]   (backward-data weights)
__global__ void conv3_flip_kernel(const bf16* __restrict__ w, int cout, int cin, bf16* __restrict__ wt,
                                  int cout_t, int cin_t) {
  // wt has "cout" = cin (padded to cout_t, a multiple of 16) and "cin" = cout
  // (padded to cin_t, a multiple of 32)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = 27LL * cout_t * cin_t;
  if (t >= total) return;
  const int tap = (int)(t / ((int64_t)cout_t * cin_t));
  const int r = (int)(t % ((int64_t)cout_t * cin_t));
  const int ci = r / cin_t, co = r % cin_t;   // wt[tap][ci][co]
  bf16 v = (bf16)0.0f;
  if (ci < cin && co < cout) v = w[((int64_t)(26 - tap) * cout + co) * cin + ci];
  wt[t] = v;
}

// dW partial for one (tap, pixel chunk): dw[co][ci] = sum_p dy[p][co] x[p+off][ci]
template <int NT>
__global__ void __launch_bounds__(kCT)
conv3_wgrad_kernel(const bf16* __restrict__ x, int cin, const bf16* __restrict__ dy, int cout, int nx, int ny,
                   int nz, int chunk, float* __restrict__ part) {
  constexpr int LD = 40;   // LDS row stride in bf16 (32 pixels + 8 pad)
  __shared__ __align__(16) bf16 sdy[64 * LD];
  __shared__ __align__(16) bf16 sx[64 * LD];
  const int64_t P = (int64_t)nx * ny * nz;
  const int tap = blockIdx.x;
  const int dz = tap / 9 - 1, dyo = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  const int64_t p0 = (int64_t)blockIdx.y * chunk;
  const int64_t p1 = (p0 + chunk < P) ? p0 + chunk : P;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int mt = wave;   // co tile of this wave
  const bool active = mt * 16 < cout;
  f32x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int64_t pb = p0; pb < p1; pb += 32) {
    // stage dy[32 px][cout] and x[32 px shifted][cin], transposed
    for (int e = threadIdx.x; e < 32 * 64; e += kCT) {
      const int pl = e / 64, c = e % 64;
      const int64_t pp = pb + pl;
      bf16 vdy = (bf16)0.0f, vx = (bf16)0.0f;
      if (pp < p1) {
        if (c < cout) vdy = dy[pp * cout + c];
        if (c < cin) {
          const int px = (int)(pp % nx), py = (int)((pp / nx) % ny), pz = (int)(pp / ((int64_t)nx * ny));
          const int qx = px + dx, qy = py + dyo, qz = pz + dz;
          if (qx >= 0 && qx < nx && qy >= 0 && qy < ny && qz >= 0 && qz < nz)
            vx = x[(((int64_t)qz * ny + qy) * nx + qx) * cin + c];
        }
      }
      sdy[c * LD + pl] = vdy;
      sx[c * LD + pl] = vx;
    }
    __syncthreads();
    if (active) {
      const bf16x8 A = *reinterpret_cast<const bf16x8*>(&sdy[(mt * 16 + m) * LD + kq * 8]);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        if (t * 16 < cin) {
          const bf16x8 B = *reinterpret_cast<const bf16x8*>(&sx[(t * 16 + m) * LD + kq * 8]);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[t], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  if (!active) return;
  float* dst = part + ((int64_t)blockIdx.y * 27 + tap) * cout * cin;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    if (t * 16 >= cin) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) dst[(mt * 16 + kq * 4 + i) * cin + t * 16 + m] = acc[t][i];
  }
}

// dw[o] = sum over chunks of part[chunk][o] in chunk order.
__global__ void sum_chunks_kernel(const float* __restrict__ part, int64_t per, int nchunk, float* __restrict__ out) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= per) return;
  float s = 0.f;
  for (int c = 0; c < nchunk; ++c) s += part[(int64_t)c * per + o];
  out[o] = s;
}

// Per-channel column sums of a (P x C) f32/bf16 matrix: partial per block.
template <typename T>
__global__ void col_sum_partial_kernel(const T* __restrict__ a, int64_t P, int C, int chunk, float* __restrict__ part) {
  const int64_t p0 = (int64_t)blockIdx.x * chunk;
  const int64_t p1 = (p0 + chunk < P) ? p0 + chunk : P;
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    float s = 0.f;
    for (int64_t p = p0; p < p1; ++p) s += (float)a[p * C + c];
    part[(int64_t)blockIdx.x * C + c] = s;
  }
}

// dz = da * (z > 0)
__global__ void relu_bwd_kernel(const bf16* __restrict__ da, const bf16* __restrict__ z, int64_t n,
                                bf16* __restrict__ dz) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  dz[t] = ((float)z[t] > 0.f) ? da[t] : (bf16)0.0f;
}

// f32 (P x cs, cs source channels) -> bf16 (P x cd), zero-padded channels
__global__ void to_bf16_kernel(const float* __restrict__ a, int64_t P, int cs, int cd, bf16* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P * cd) return;
  const int64_t p = t / cd;
  const int c = (int)(t % cd);
  out[t] = (bf16)((c < cs) ? a[p * cs + c] : 0.f);
}

int conv_check(int cin, int cout, int nx, int ny, int nz) {
  FEM_REQUIRE(cin > 0 && cin % 32 == 0 && cin <= 64, FEM_E_INVALID, "cin must be 32 or 64 (padded), got %d", cin);
  FEM_REQUIRE(cout > 0 && cout % 16 == 0 && cout <= 64, FEM_E_INVALID, "cout must be a multiple of 16 <= 64");
  FEM_REQUIRE(nx > 0 && ny > 0 && nz > 0, FEM_E_INVALID, "bad grid");
  return FEM_OK;
}

}  // namespace
}  // namespace fem

using namespace fem;

extern "C" int fem_meshconv_fwd_bf16(const void* x, int cin, const void* w, const float* bias, int cout, int nx,
                                     int ny, int nz, void* y_act, void* y_pre, float* y_f32, int relu,
                                     void* stream) {
  int rc = conv_check(cin, cout, nx, ny, nz);
  if (rc) return rc;
  FEM_REQUIRE(x && w && (y_act || y_pre || y_f32), FEM_E_INVALID, "null pointer");
  const dim3 grid((unsigned)ceil_div(nx, kTX), (unsigned)ceil_div(ny, kTY), (unsigned)ceil_div(nz, kTZ));
  FEM_REQUIRE(grid.y <= 65535 && grid.z <= 65535, FEM_E_LIMIT, "grid too large for the tiled convolution");
  hipStream_t st = as_stream(stream);
  const bf16* xb = (const bf16*)x;
  const bf16* wb = (const bf16*)w;
  bf16 *ya = (bf16*)y_act, *yp = (bf16*)y_pre;
  switch (cout / 16) {
    case 1: conv3_fwd_tiled_kernel<1><<<grid, kCT, 0, st>>>(xb, cin, wb, bias, cout, nx, ny, nz, ya, yp, y_f32, relu); break;
    case 2: conv3_fwd_tiled_kernel<2><<<grid, kCT, 0, st>>>(xb, cin, wb, bias, cout, nx, ny, nz, ya, yp, y_f32, relu); break;
    case 3: conv3_fwd_tiled_kernel<3><<<grid, kCT, 0, st>>>(xb, cin, wb, bias, cout, nx, ny, nz, ya, yp, y_f32, relu); break;
    default: conv3_fwd_tiled_kernel<4><<<grid, kCT, 0, st>>>(xb, cin, wb, bias, cout, nx, ny, nz, ya, yp, y_f32, relu); break;
  }
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_meshconv_flip_bf16(const void* w, int cout, int cin, void* wt, int cout_t, int cin_t,
                                      void* stream) {
  FEM_REQUIRE(w && wt && cout_t % 16 == 0 && cin_t % 32 == 0 && cout_t >= cin && cin_t >= cout, FEM_E_INVALID,
              "bad flip arguments");
  const int64_t total = 27LL * cout_t * cin_t;
  conv3_flip_kernel<<<ceil_div(total, 256), 256, 0, as_stream(stream)>>>((const bf16*)w, cout, cin, (bf16*)wt,
                                                                        cout_t, cin_t);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int64_t fem_meshconv_wgrad_work_bytes(int cin, int cout, int nx, int ny, int nz, int chunk) {
  const int64_t P = (int64_t)nx * ny * nz;
  return 4LL * ceil_div(P, chunk) * 27 * cin * cout + 256;
}

extern "C" int fem_meshconv_wgrad_bf16(const void* x, int cin, const void* dy, int cout, int nx, int ny, int nz,
                                       int chunk, float* dw, void* work, int64_t work_bytes, void* stream) {
  int rc = conv_check(cin, cout, nx, ny, nz);
  if (rc) return rc;
  FEM_REQUIRE(chunk > 0 && chunk % 32 == 0, FEM_E_INVALID, "chunk must be a positive multiple of 32");
  FEM_REQUIRE(x && dy && dw && work, FEM_E_INVALID, "null pointer");
  FEM_REQUIRE(work_bytes >= fem_meshconv_wgrad_work_bytes(cin, cout, nx, ny, nz, chunk), FEM_E_INVALID,
              "work too small");
  const int64_t P = (int64_t)nx * ny * nz;
  const int nchunk = (int)ceil_div(P, chunk);
  hipStream_t st = as_stream(stream);
  float* part = reinterpret_cast<float*>(work);
  const dim3 grid(27, nchunk);
  conv3_wgrad_kernel<4><<<grid, kCT, 0, st>>>((const bf16*)x, cin, (const bf16*)dy, cout, nx, ny, nz, chunk, part);
  const int64_t per = 27LL * cout * cin;
  sum_chunks_kernel<<<ceil_div(per, 256), 256, 0, st>>>(part, per, nchunk, dw);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int64_t fem_colsum_work_bytes(int64_t P, int C, int chunk) { return 4LL * ceil_div(P, chunk) * C + 256; }

extern "C" int fem_colsum(const void* a, int is_bf16, int64_t P, int C, int chunk, float* out, void* work,
                          int64_t work_bytes, void* stream) {
  FEM_REQUIRE(a && out && work && P > 0 && C > 0 && chunk > 0, FEM_E_INVALID, "bad arguments");
  FEM_REQUIRE(work_bytes >= fem_colsum_work_bytes(P, C, chunk), FEM_E_INVALID, "work too small");
  const int nchunk = (int)ceil_div(P, chunk);
  hipStream_t st = as_stream(stream);
  float* part = reinterpret_cast<float*>(work);
  if (is_bf16)
    col_sum_partial_kernel<bf16><<<nchunk, 64, 0, st>>>((const bf16*)a, P, C, chunk, part);
  else
    col_sum_partial_kernel<float><<<nchunk, 64, 0, st>>>((const float*)a, P, C, chunk, part);
  sum_chunks_kernel<<<ceil_div(C, 256), 256, 0, st>>>(part, C, nchunk, out);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_relu_bwd_bf16(const void* da, const void* z, int64_t n, void* dz, void* stream) {
  FEM_REQUIRE(da && z && dz && n >= 0, FEM_E_INVALID, "bad arguments");
  if (n == 0) return FEM_OK;
  relu_bwd_kernel<<<ceil_div(n, 256), 256, 0, as_stream(stream)>>>((const bf16*)da, (const bf16*)z, n, (bf16*)dz);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_to_bf16(const float* a, int64_t P, int cs, int cd, void* out, void* stream) {
  FEM_REQUIRE(a && out && P >= 0 && cs > 0 && cd >= cs, FEM_E_INVALID, "bad arguments");
  if (P == 0) return FEM_OK;
  to_bf16_kernel<<<ceil_div(P * cd, 256), 256, 0, as_stream(stream)>>>(a, P, cs, cd, (bf16*)out);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_meshconv_fwd_c4_bf16(const void* x4, const void* wp, const float* bias, int cout, int nx, int ny,
                                        int nz, void* y_act, void* y_pre, int relu, void* stream) {
  FEM_REQUIRE(cout > 0 && cout % 16 == 0 && cout <= 64, FEM_E_INVALID, "cout must be a multiple of 16 <= 64");
  FEM_REQUIRE(nx > 0 && ny > 0 && nz > 0, FEM_E_INVALID, "bad grid");
  FEM_REQUIRE(x4 && wp && (y_act || y_pre), FEM_E_INVALID, "null pointer");
  FEM_REQUIRE((reinterpret_cast<uintptr_t>(x4) & 7) == 0 && (reinterpret_cast<uintptr_t>(wp) & 15) == 0,
              FEM_E_INVALID, "x4 must be 8-byte and wp 16-byte aligned");
  const dim3 grid((unsigned)ceil_div(nx, kTX), (unsigned)ceil_div(ny, kTY), (unsigned)ceil_div(nz, kTZ));
  FEM_REQUIRE(grid.y <= 65535 && grid.z <= 65535, FEM_E_LIMIT, "grid too large for the tiled convolution");
  hipStream_t st = as_stream(stream);
  const bf16* xb = (const bf16*)x4;
  const bf16* wb = (const bf16*)wp;
  bf16 *ya = (bf16*)y_act, *yp = (bf16*)y_pre;
  switch (cout / 16) {
    case 1: conv3_fwd_c4_kernel<1><<<grid, kCT, 0, st>>>(xb, wb, bias, cout, nx, ny, nz, ya, yp, relu); break;
    case 2: conv3_fwd_c4_kernel<2><<<grid, kCT, 0, st>>>(xb, wb, bias, cout, nx, ny, nz, ya, yp, relu); break;
    case 3: conv3_fwd_c4_kernel<3><<<grid, kCT, 0, st>>>(xb, wb, bias, cout, nx, ny, nz, ya, yp, relu); break;
    default: conv3_fwd_c4_kernel<4><<<grid, kCT, 0, st>>>(xb, wb, bias, cout, nx, ny, nz, ya, yp, relu); break;
  }
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_meshconv_pack_c4_bf16(const void* w, int cin_w, int cout, void* wp, void* stream) {
  FEM_REQUIRE(w && wp && cin_w >= 4 && cout > 0 && cout % 16 == 0 && cout <= 64, FEM_E_INVALID,
              "bad packing arguments");
  const int n = 4 * cout * 32;
  pack_c4_kernel<<<ceil_div(n, 256), 256, 0, as_stream(stream)>>>((const bf16*)w, cin_w, cout, (bf16*)wp);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_meshconv_features_c4(const double* r, int64_t P, const double* scale, const float* efeat,
                                        void* x4, void* stream) {
  FEM_REQUIRE(P >= 0, FEM_E_INVALID, "negative node count");
  if (P == 0) return FEM_OK;
  FEM_REQUIRE(r && scale && efeat && x4, FEM_E_INVALID, "null pointer");
  features_c4_kernel<<<ceil_div(4 * P, 256), 256, 0, as_stream(stream)>>>(r, P, scale, efeat, (bf16*)x4);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_meshconv_correction(const float* y, int ld, int64_t P, const double* scale, double e_ref,
                                       const double* mask, double* c, void* stream) {
  FEM_REQUIRE(P >= 0 && ld >= 3, FEM_E_INVALID, "bad size");
  if (P == 0) return FEM_OK;
  FEM_REQUIRE(y && scale && c, FEM_E_INVALID, "null pointer");
  correction_kernel<<<ceil_div(3 * P, 256), 256, 0, as_stream(stream)>>>(y, ld, P, scale, e_ref, mask, c);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}
