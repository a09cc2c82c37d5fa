// hex8 stiffness / mass assembly into the node-pair pattern, row-owner and
// atomic-free.
//
// The reference (FEM_model_csr.py:150-165, 246-257; FEM_model_bsr.py:163-175)
// computes an E x 8 x 8 x 3 x 3 element intermediate and scatters it with
// atomic += after a linear slot search.  Here one work-group owns a group of
// node rows (a spatially compact set built once from the pattern):
//   phase 1: one thread per (element, Gauss point) of the elements touching
//            the rows computes J and stages in LDS the physical gradients
//            a[e][k][g] = DNr_g[k] J_g^-1 and w[e][g] = E_e * detJ_g;
//   phase 2: one lane per task.  A task holds at most two element
//            contributions of one output block (row i, column node c): the
//            3x3 Gram block sum_g w_g a_gj (x) a_gk with j, k the local
//            indices of i and c in the element.  The T tasks of a block are
//            T adjacent, T-aligned lanes and are summed by a fixed butterfly
//            of lane shuffles, so every block is a sum in a fixed order --
//            bit-identical run to run, no atomics, no LDS accumulator.  The
//            head lane writes K_block = lhat*G + mhat*G^T + mhat*tr(G)*I.
//            Blocks with a single contribution are paired two per lane.
// The closed form is B_j^T D B_k with D = E*[lhat 1 1^T + 2 mhat I on the
// normal block, mhat I on the shear block] (lhat = nu/((1+nu)(1-2nu)),
// mhat = 1/(2(1+nu))): K is linear in G and E is folded into w, so the
// combination is applied once per block after the sum over elements.
// The task lists are symbolic metadata built once per mesh (groups.py).
#include "common.hpp"

namespace fem {
namespace {

constexpr int kAsmThreads = 512;
// (e, k) row stride of a[] in elements of T.  FP64: 24 used + 2 pad; 26
// doubles = 208 B keeps every row 16-byte aligned for 128-bit LDS reads and
// spreads rows over 16 distinct groups of 4 banks.  FP32: 24 + 1, an odd
// number of 4-byte banks between rows.
template <typename T>
__host__ __device__ constexpr int a_stride() { return sizeof(T) == 8 ? 26 : 25; }
constexpr int kNone = 0x3FFF;

struct AsmLds {
  int a_off, w_off, wm_off, n_off, sm_off, co_off, pre_off, total;
};

__host__ __device__ inline int align16(int o) { return (o + 15) & ~15; }

template <typename T>
__host__ __device__ inline AsmLds asm_lds_layout(int eg, int rg, bool want_k, bool want_m) {
  AsmLds L;
  int o = 0;
  L.a_off = o;
  if (want_k) o += eg * 8 * a_stride<T>() * (int)sizeof(T);
  o = align16(o);
  L.w_off = o;
  if (want_k) o += eg * 8 * (int)sizeof(T);
  o = align16(o);
  L.wm_off = o;
  if (want_m) o += eg * 8 * (int)sizeof(double);
  o = align16(o);
  L.n_off = o;
  if (want_m) o += 64 * (int)sizeof(double);
  L.sm_off = o;
  o += rg * (int)sizeof(int32_t);
  L.co_off = o;
  o += rg * (int)sizeof(int32_t);
  L.pre_off = o;
  o += rg * (int)sizeof(int32_t);
  L.total = align16(o);
  return L;
}

// dN_k/dxi at a Gauss point from its factors f_a = 1 + s_ka * sg_a * GS.
__device__ __forceinline__ void dn_factored(int k, const double fp[3], const double fm[3], double dn[3]) {
  const double f0 = node_sign(k, 0) > 0 ? fp[0] : fm[0];
  const double f1 = node_sign(k, 1) > 0 ? fp[1] : fm[1];
  const double f2 = node_sign(k, 2) > 0 ? fp[2] : fm[2];
  dn[0] = 0.125 * node_sign(k, 0) * (f1 * f2);
  dn[1] = 0.125 * node_sign(k, 1) * (f0 * f2);
  dn[2] = 0.125 * node_sign(k, 2) * (f0 * f1);
}

// 8 dN_k/dxi: the same without the 1/8.  With it J' = 8 J, whose cofactor
// inverse is J^-1 / 8 and whose determinant is 512 det J, all exactly (powers
// of two), so dN8 J'^-1 = dN J^-1 bit for bit and no scaling instruction runs.
__device__ __forceinline__ void dn8_factored(int k, const double fp[3], const double fm[3], double dn[3]) {
  const double f0 = node_sign(k, 0) > 0 ? fp[0] : fm[0];
  const double f1 = node_sign(k, 1) > 0 ? fp[1] : fm[1];
  const double f2 = node_sign(k, 2) > 0 ? fp[2] : fm[2];
  dn[0] = node_sign(k, 0) * (f1 * f2);
  dn[1] = node_sign(k, 1) * (f0 * f2);
  dn[2] = node_sign(k, 2) * (f0 * f1);
}
constexpr double kInv512 = 1.0 / 512.0;

template <typename T>
__device__ __forceinline__ void gram(const T* __restrict__ s_a, const T* __restrict__ s_w, int c, T G[9]) {
  // G += sum_g w_g a_gj (x) a_gk for contribution c (scaled by 0 when absent)
  const bool has = c != kNone;
  const int cc = has ? c : 0;
  const int le = cc >> 6, j = (cc >> 3) & 7, k = cc & 7;
  const T* aj = s_a + (le * 8 + j) * a_stride<T>();
  const T* ak = s_a + (le * 8 + k) * a_stride<T>();
  const T* w = s_w + le * 8;
  if constexpr (sizeof(T) == 8) {
    // two Gauss points per step: 6 doubles of each row = three 128-bit reads
#pragma unroll 2
    for (int gp = 0; gp < 4; ++gp) {
      const double2* pj = reinterpret_cast<const double2*>(aj + 6 * gp);
      const double2* pk = reinterpret_cast<const double2*>(ak + 6 * gp);
      const double2 ww = reinterpret_cast<const double2*>(w)[gp];
      const double2 j0 = pj[0], j1 = pj[1], j2 = pj[2];
      const double2 k0 = pk[0], k1 = pk[1], k2 = pk[2];
      const double wa = has ? ww.x : 0.0, wb = has ? ww.y : 0.0;
      const double ja[3] = {j0.x, j0.y, j1.x}, jb[3] = {j1.y, j2.x, j2.y};
      const double ka[3] = {k0.x, k0.y, k1.x}, kb[3] = {k1.y, k2.x, k2.y};
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        const double u = wa * ja[p], v = wb * jb[p];
#pragma unroll
        for (int q = 0; q < 3; ++q) G[3 * p + q] = fma(v, kb[q], fma(u, ka[q], G[3 * p + q]));
      }
    }
  } else {
#pragma unroll 2
    for (int g = 0; g < 8; ++g) {
      const T wg = has ? w[g] : T(0);
      const T u0 = wg * aj[3 * g], u1 = wg * aj[3 * g + 1], u2 = wg * aj[3 * g + 2];
      const T b0 = ak[3 * g], b1 = ak[3 * g + 1], b2 = ak[3 * g + 2];
      G[0] = fma(u0, b0, G[0]); G[1] = fma(u0, b1, G[1]); G[2] = fma(u0, b2, G[2]);
      G[3] = fma(u1, b0, G[3]); G[4] = fma(u1, b1, G[4]); G[5] = fma(u1, b2, G[5]);
      G[6] = fma(u2, b0, G[6]); G[7] = fma(u2, b1, G[7]); G[8] = fma(u2, b2, G[8]);
    }
  }
}

// Lane exchange inside a row of 16 lanes by DPP (VALU, no LDS traffic).
// CTRL 0xB1 / 0x4E: quad_perm [1,0,3,2] / [2,3,0,1] (partner lane ^1 / ^2);
// 0x141: row_half_mirror (lane i <-> 7-i of its 8); 0x140: row_mirror
// (i <-> 15-i of its 16).  After ^1 and ^2 every lane of a quad holds the
// quad's sum, so pairing with the mirrored lane of the other quad (other
// half-row) completes the sum just as ^4 (^8) would.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over the T adjacent, T-aligned lanes of each block (T <= 32); every
// lane of the block ends with the same total.  Steps no lane of the wave
// needs are skipped (wave-uniform test).
template <typename T, int N>
__device__ __forceinline__ void block_allsum(T (&G)[N], int Tn) {
  if (__any(Tn > 1)) {
#pragma unroll
    for (int q = 0; q < N; ++q) { const T o = dpp_move<0xB1>(G[q]); if (Tn > 1) G[q] += o; }
  }
  if (__any(Tn > 2)) {
#pragma unroll
    for (int q = 0; q < N; ++q) { const T o = dpp_move<0x4E>(G[q]); if (Tn > 2) G[q] += o; }
  }
  if (__any(Tn > 4)) {
#pragma unroll
    for (int q = 0; q < N; ++q) { const T o = dpp_move<0x141>(G[q]); if (Tn > 4) G[q] += o; }
  }
  if (__any(Tn > 8)) {
#pragma unroll
    for (int q = 0; q < N; ++q) { const T o = dpp_move<0x140>(G[q]); if (Tn > 8) G[q] += o; }
  }
  if (__any(Tn > 16)) {
#pragma unroll
    for (int q = 0; q < N; ++q) { const T o = __shfl_xor(G[q], 16, 64); if (Tn > 16) G[q] += o; }
  }
}

__device__ __forceinline__ double mass_term(const double* __restrict__ s_wm, const double* __restrict__ s_N, int c) {
  if (c == kNone) return 0.0;
  const int le = c >> 6, j = (c >> 3) & 7, k = c & 7;
  double ms = 0.0;
#pragma unroll
  for (int g = 0; g < 8; ++g) ms = fma(s_wm[le * 8 + g] * s_N[g * 8 + j], s_N[g * 8 + k], ms);
  return ms;
}

// Contraction off and the one fused multiply-add explicit, so every call site
// (direct stores of either layout, the staged copy) rounds the same way.
template <typename T>
__device__ __forceinline__ void k_block(const T G[9], T lh, T mh, T v[9]) {
#pragma clang fp contract(off)
  const T tr = mh * (G[0] + G[4] + G[8]);
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const T x = fma(lh, G[3 * p + q], mh * G[3 * q + p]);
      v[3 * p + q] = p == q ? x + tr : x;
    }
}

template <int LAYOUT>
__device__ __forceinline__ int64_t k_addr(int64_t sm, int co, int m, int idx) {
  if (LAYOUT == FEM_LAYOUT_BSR) return 9 * (sm + m) + idx;
  const int p = (idx * 11) >> 5, q = idx - 3 * p;   // idx / 3, idx % 3 for idx < 9
  return 9 * sm + 3 * (int64_t)p * co + 3 * m + q;
}

// All 9 values of a block from one lane.
template <typename T, int LAYOUT>
__device__ __forceinline__ void store_block(T* __restrict__ k_vals, const int32_t* s_sm, const int32_t* s_co,
                                            int out, const T G[9], T lh, T mh) {
  const int r = out >> 8, m = out & 255;
  const int64_t sm = s_sm[r];
  const int co = s_co[r];
  T v[9];
  k_block(G, lh, mh, v);
#pragma unroll
  for (int idx = 0; idx < 9; ++idx) k_vals[k_addr<LAYOUT>(sm, co, m, idx)] = v[idx];
}

// One K block into the staging copy of the group's rows (row r's blocks from
// 9 * s_pre[r] on, in column order, as they lie in the BSR values).
template <typename T>
__device__ __forceinline__ void stage_block(T* __restrict__ st, const int32_t* s_pre, int out, const T G[9], T lh,
                                            T mh) {
  T v[9];
  k_block(G, lh, mh, v);
  T* dst = st + 9 * (s_pre[out >> 8] + (out & 255));
#pragma unroll
  for (int idx = 0; idx < 9; ++idx) dst[idx] = v[idx];
}

// 4 waves per SIMD (<= 128 VGPRs): two work-groups of 512 threads per CU.
// stage != 0 (BSR, set by the host when every group's blocks fit in the
// gradient area): a single-pass group stages its finished blocks in LDS over
// the gradients and copies each row's contiguous BSR run out with 32 lanes,
// instead of one 72-byte block per lane (one store instruction then touches
// about 4 cache lines instead of about 36).
template <typename T, bool WANT_K, bool WANT_M, int LAYOUT>
__global__ void __launch_bounds__(kAsmThreads, 4)
hex8_assemble_kernel(fem_mesh m, const double* __restrict__ young, double young_scalar,
                     double lhat, double mhat, double rho, T* __restrict__ k_vals,
                     double* __restrict__ m_blocks, int eg, int rg, int stage) {
  extern __shared__ __align__(16) unsigned char smem[];
  const AsmLds L = asm_lds_layout<T>(eg, rg, WANT_K, WANT_M);
  T* s_a = reinterpret_cast<T*>(smem + L.a_off);
  T* s_w = reinterpret_cast<T*>(smem + L.w_off);
  double* s_wm = reinterpret_cast<double*>(smem + L.wm_off);
  double* s_N = reinterpret_cast<double*>(smem + L.n_off);
  int32_t* s_sm = reinterpret_cast<int32_t*>(smem + L.sm_off);
  int32_t* s_co = reinterpret_cast<int32_t*>(smem + L.co_off);
  int32_t* s_pre = reinterpret_cast<int32_t*>(smem + L.pre_off);

  const int grp = xcd_chunked(blockIdx.x, gridDim.x);
  const int e0 = m.grp_elem_ptr[grp];
  const int ne = m.grp_elem_ptr[grp + 1] - e0;
  const int r0 = m.grp_row_ptr[grp];
  const int nr = m.grp_row_ptr[grp + 1] - r0;
  const int t0 = m.grp_task_ptr[grp];
  const int nt = m.grp_task_ptr[grp + 1] - t0;
  const int tid = threadIdx.x;

  // Every load that depends only on the group is issued here, before any
  // wait: this lane's first task, its element's node list and modulus, and
  // (for the first nr lanes) the row pointers.
  const int2 none_task = make_int2((int)0xFFFF0000u, kNone | (kNone << 14));
  int2 task = tid < nt ? reinterpret_cast<const int2*>(m.tasks)[t0 + tid] : none_task;
  const bool p1 = tid < ne * 8;            // ne <= 64 (checked on the host)
  const int le = tid >> 3, g = tid & 7;
  int4 na = make_int4(0, 0, 0, 0), nb = na;
  double Ee = young_scalar;
  if (p1) {
    const int4* en = reinterpret_cast<const int4*>(m.grp_enodes + (int64_t)(e0 + le) * 8);
    na = en[0];
    nb = en[1];
    if (young) Ee = young[m.grp_elems[e0 + le]];
  }
  if (tid < nr) {
    const int32_t i = m.grp_rows[r0 + tid];
    s_sm[tid] = m.sm_node[i];
    s_co[tid] = m.co_node[i];
  }
  if (WANT_M && tid < 64) s_N[tid] = nshape(tid >> 3, tid & 7);   // s_N[g*8 + k]

  // ---- phase 1: gradients and weights at each (element, Gauss point)
  if (p1) {
    const int32_t nodes[8] = {na.x, na.y, na.z, na.w, nb.x, nb.y, nb.z, nb.w};
    // 1 +- GS along each axis at this Gauss point: dN_k/dxi_d = s_kd/8 * prod_{a!=d} f_a(k)
    double fp[3], fm[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      fp[a] = 1.0 + gauss_sign(g, a) * kGS;
      fm[a] = 1.0 - gauss_sign(g, a) * kGS;
    }
    const int64_t n0 = nodes[0];
    const double x00 = m.coords[n0 * 3], x01 = m.coords[n0 * 3 + 1], x02 = m.coords[n0 * 3 + 2];
    double J[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) J[q] = 0.0;
#pragma unroll
    for (int k = 1; k < 8; ++k) {   // 8 J = sum_k (x_k - x_0) (x) 8 dN_k (the x_0 term sums to zero)
      const int64_t nk = nodes[k];
      const double xr[3] = {m.coords[nk * 3] - x00, m.coords[nk * 3 + 1] - x01, m.coords[nk * 3 + 2] - x02};
      double dn[3];
      dn8_factored(k, fp, fm, dn);
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) J[a * 3 + d] = fma(xr[a], dn[d], J[a * 3 + d]);
    }
    if (WANT_K) {
      double Ji[9];
      const double det = inv3(J, Ji) * kInv512;     // det J
      s_w[le * 8 + g] = (T)(Ee * det);
      if (WANT_M) s_wm[le * 8 + g] = rho * det;
      // recompute the factors rather than keep 24 derivatives live across the inverse
#pragma unroll
      for (int a = 0; a < 3; ++a) asm volatile("" : "+v"(fp[a]), "+v"(fm[a]));
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        double dn[3];
        dn8_factored(k, fp, fm, dn);
        T* dst = s_a + (le * 8 + k) * a_stride<T>() + 3 * g;
#pragma unroll
        for (int a = 0; a < 3; ++a) dst[a] = (T)(dn[0] * Ji[0 * 3 + a] + dn[1] * Ji[1 * 3 + a] + dn[2] * Ji[2 * 3 + a]);
      }
    } else if (WANT_M) {
      s_wm[le * 8 + g] = rho * (det3(J) * kInv512);
    }
  }
  __syncthreads();

  // row offsets of the staging copy: exclusive prefix of the rows' block counts
  const bool staged = WANT_K && LAYOUT == FEM_LAYOUT_BSR && stage != 0 && nt <= kAsmThreads;   // uniform
  if (staged && tid < 64) {
    const int own = tid < nr ? s_co[tid] : 0;
    int v = own;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(v, o, 64);
      if (tid >= o) v += u;
    }
    if (tid < nr) s_pre[tid] = v - own;
  }

  // ---- phase 2: one lane per task, T-lane butterfly per output block
  const T lh = (T)lhat, mh = (T)mhat;
  const int lane = tid & 63;
  for (int base = 0; base < nt; base += kAsmThreads) {   // uniform trip count: all lanes shuffle
    const int t = base + tid;
    const bool valid = t < nt;
    if (base > 0) task = valid ? reinterpret_cast<const int2*>(m.tasks)[t0 + t] : none_task;
    const int outA = task.x & 0xFFFF, outB = (task.x >> 16) & 0xFFFF;
    const uint32_t w2 = (uint32_t)task.y;
    const int c1 = (int)(w2 & 0x3FFF), c2 = (int)((w2 >> 14) & 0x3FFF);
    const bool dual = (w2 >> 31) & 1u;
    const int lt = (int)((w2 >> 28) & 7u);
    const int Tn = dual ? 1 : (1 << lt);
    const bool has_b = dual && outB != 0xFFFF;
    if (WANT_K) {
      T G1[9], G2[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) G1[q] = G2[q] = T(0);
      gram<T>(s_a, s_w, c1, G1);
      gram<T>(s_a, s_w, c2, G2);
      if (dual) {   // second output of a dual task: no butterfly, store now
        if (!staged && valid && has_b) store_block<T, LAYOUT>(k_vals, s_sm, s_co, outB, G2, lh, mh);
      } else {
#pragma unroll
        for (int q = 0; q < 9; ++q) G1[q] += G2[q];
      }
      block_allsum(G1, Tn);
      const bool head = valid && (lane & (Tn - 1)) == 0;
      if (staged) {
        __syncthreads();   // every wave is done with the gradients: their area becomes the staging copy
        T* st = reinterpret_cast<T*>(smem + L.a_off);
        if (head) stage_block<T>(st, s_pre, outA, G1, lh, mh);
        if (valid && has_b) stage_block<T>(st, s_pre, outB, G2, lh, mh);
        __syncthreads();
        for (int r = tid >> 5; r < nr; r += kAsmThreads / 32) {
          const int n9 = 9 * s_co[r];
          T* __restrict__ dst = k_vals + 9 * (int64_t)s_sm[r];
          const T* src = st + 9 * s_pre[r];
          for (int d = tid & 31; d < n9; d += 32) dst[d] = src[d];
        }
      } else if (head) {
        store_block<T, LAYOUT>(k_vals, s_sm, s_co, outA, G1, lh, mh);
      }
    }
    if (WANT_M) {
      double m1[1] = {mass_term(s_wm, s_N, c1)};
      const double m2 = mass_term(s_wm, s_N, c2);
      if (!dual) m1[0] += m2;
      block_allsum(m1, Tn);
      if (valid && (lane & (Tn - 1)) == 0) {
        m_blocks[(int64_t)s_sm[outA >> 8] + (outA & 255)] = m1[0];
        if (has_b) m_blocks[(int64_t)s_sm[outB >> 8] + (outB & 255)] = m2;
      }
    }
  }
}

template <typename T, bool K, bool M, int LAYOUT>
int launch_asm(const fem_mesh& m, const double* young, double ys, double nu, double rho,
               void* k_vals, double* m_blocks, hipStream_t st) {
  const int eg = m.max_group_elems, rg = m.max_group_rows;
  const AsmLds L = asm_lds_layout<T>(eg, rg, K, M);
  FEM_REQUIRE(L.total <= 160 * 1024, FEM_E_LIMIT,
              "row group needs %d B of LDS (> 160 KB): use smaller row groups", L.total);
  auto kern = hex8_assemble_kernel<T, K, M, LAYOUT>;
  FEM_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, L.total));
  const double lhat = nu / ((1 + nu) * (1 - 2 * nu));
  const double mhat = 1.0 / (2 * (1 + nu));
  // staging needs the group's rows in one wave (prefix scan) and their blocks inside the gradient area
  const int stage = K && LAYOUT == FEM_LAYOUT_BSR && rg <= 64 &&
                    (int64_t)m.max_group_slots * 9 * (int64_t)sizeof(T) <= (int64_t)(L.w_off - L.a_off);
  hipLaunchKernelGGL(kern, dim3((unsigned)m.num_groups), dim3(kAsmThreads), L.total, st, m, young,
                     ys, lhat, mhat, rho, reinterpret_cast<T*>(k_vals), m_blocks, eg, rg, stage);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

// ---------------------------------------------------------------------------
// CSR index arrays and value expansion (FEM_model_csr.py:163, 168-173, 189-201)
// ---------------------------------------------------------------------------
__global__ void csr_index_K_kernel(fem_mesh m, int32_t* __restrict__ row_k, int32_t* __restrict__ col_k) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < m.num_node) {
    const int32_t co = m.co_node[s], sm = m.sm_node[s];
    for (int p = 0; p < 3; ++p) row_k[3 * s + 1 + p] = co * 3 * (p + 1) + sm * 9;
    if (s == 0) row_k[0] = 0;
  }
  if (s >= m.nnzb) return;
  // row of slot s: last i with sm_node[i] <= s (binary search)
  int64_t lo = 0, hi = m.num_node - 1;
  while (lo < hi) {
    int64_t mid = (lo + hi + 1) >> 1;
    if (m.sm_node[mid] <= s) lo = mid; else hi = mid - 1;
  }
  const int64_t i = lo;
  const int64_t sm = m.sm_node[i], co = m.co_node[i], mloc = s - sm;
  const int32_t c = m.re_node[s];
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q) col_k[9 * sm + 3 * p * co + 3 * mloc + q] = 3 * c + q;
}

__global__ void csr_index_M_kernel(fem_mesh m, int32_t* __restrict__ row_m, int32_t* __restrict__ col_m,
                                   const double* __restrict__ m_blocks, double* __restrict__ val_m) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row_m && s < m.num_node) {
    const int32_t co = m.co_node[s], sm = m.sm_node[s];
    for (int p = 0; p < 3; ++p) row_m[3 * s + 1 + p] = co * (p + 1) + sm * 3;
    if (s == 0) row_m[0] = 0;
  }
  if (s >= m.nnzb) return;
  int64_t lo = 0, hi = m.num_node - 1;
  while (lo < hi) {
    int64_t mid = (lo + hi + 1) >> 1;
    if (m.sm_node[mid] <= s) lo = mid; else hi = mid - 1;
  }
  const int64_t i = lo;
  const int64_t sm = m.sm_node[i], co = m.co_node[i], mloc = s - sm;
  const int32_t c = m.re_node[s];
  for (int p = 0; p < 3; ++p) {
    const int64_t idx = 3 * sm + p * co + mloc;
    if (col_m) col_m[idx] = 3 * c + p;
    if (val_m) val_m[idx] = m_blocks[s];
  }
}

// ---------------------------------------------------------------------------
// Spin softening (no counterpart in the reference; Readme.md:189 lists it as
// missing): the centrifugal load rho (|w|^2 x - (w.x) w) grows with the
// displacement, a negative stiffness K_jk -= m_jk (|w|^2 I - w w^T), m_jk the
// consistent-mass scalar of the node pair.  Thread per block row.
// ---------------------------------------------------------------------------
template <int LAYOUT>
__global__ void spin_kernel(fem_mesh m, const double* __restrict__ mb, double w0, double w1, double w2,
                            double* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.num_node) return;
  const int64_t sm = m.sm_node[i];
  const int co = m.co_node[i];
  const double w[3] = {w0, w1, w2};
  const double ww = w0 * w0 + w1 * w1 + w2 * w2;
  for (int k = 0; k < co; ++k) {
    const double mk = mb[sm + k];
#pragma unroll
    for (int idx = 0; idx < 9; ++idx) {
      const int p = idx / 3, q = idx % 3;
      vals[k_addr<LAYOUT>(sm, co, k, idx)] -= mk * ((p == q ? ww : 0.0) - w[p] * w[q]);
    }
  }
}

// ---------------------------------------------------------------------------
// Penalty (FEM_model_csr.py:233-242 / FEM_model_bsr.py:248-259): SET diagonal.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void penalty_kernel(fem_mesh m, const int32_t* __restrict__ bnode,
                               const double* __restrict__ bflag, int64_t nb, T* __restrict__ vals,
                               int layout, double value) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int32_t i = bnode[b];
  if (i < 0 || i >= m.num_node) return;
  const int32_t sm = m.sm_node[i], co = m.co_node[i];
  int k = -1;
  for (int q = 0; q < co; ++q)
    if (m.re_node[sm + q] == i) { k = q; break; }
  if (k < 0) return;
  for (int j = 0; j < 3; ++j) {
    if (bflag[3 * b + j] == -1.0) continue;
    int64_t idx = (layout == FEM_LAYOUT_BSR) ? 9 * ((int64_t)sm + k) + 4 * j
                                             : 9 * (int64_t)sm + 3 * (int64_t)j * co + 3 * k + j;
    vals[idx] = (T)value;
  }
}

// ---------------------------------------------------------------------------
// Large-displacement matrices (UnitK, FEM_model_bsr.py:476-598): the tangent
// KT and the matrix KM of the residual KM U - F, at the displacement u.
// Closed form of the reference's B / BN products, per Gauss point, with
// F = I + du/dx, Fh = I + du/dx / 2, A = F a, H = Fh a (a: physical shape
// gradients), lambda = E lhat, mu = E mhat:
//   KT_jk += detJ [lambda A_j A_k^T + mu A_k A_j^T + mu (a_j.a_k) F F^T] + detJ (a_j^T s a_k) I
//   KM_jk += detJ [lambda A_j H_k^T + mu A_k H_j^T + mu (a_j.a_k) F Fh^T]
// where F Fh^T = (F + F F^T) / 2 and s is the element stress accumulated
// over the Gauss points visited so far (the reference sums S = lambda tr(E) I
// + 2 mu E, E the Green strain, into one element field inside its Gauss
// loop), taken as an inclusive scan over the element's 8 Gauss lanes.  Same
// row groups / task lists / lane butterfly as the linear kernel; FP64 only.
//
// MODE 1 (stress stiffening of a linear prestress, no counterpart in the
// reference, whose Readme.md:189 lists it as missing): F = I and s = the
// Gauss point's own linear stress lambda tr(eps) I + 2 mu eps, eps =
// sym(du/dx), so KT = K_L + K_sigma(u) with K_sigma_jk = sum_g detJ
// (a_j^T s_g a_k) I; only KT is written.
// ---------------------------------------------------------------------------
constexpr int kKnThreads = 256;
constexpr int kKnG = 24;   // per (e, g): wl, wm, F[9], FFt[6], st[6] (xx yy zz xy xz yz), pad

struct KnLds {
  int a_off, g_off, sm_off, co_off, total;
};

__host__ __device__ inline KnLds kn_lds_layout(int eg, int rg) {
  KnLds L;
  int o = 0;
  L.a_off = o;
  o += eg * 8 * a_stride<double>() * 8;
  L.g_off = o;
  o += eg * 8 * kKnG * 8;
  L.sm_off = o;
  o += rg * 4;
  L.co_off = o;
  o += rg * 4;
  L.total = align16(o);
  return L;
}

__device__ __forceinline__ void kn_contrib(const double* __restrict__ s_a, const double* __restrict__ s_g, int c,
                                           double KT[9], double KM[9]) {
  const bool has = c != kNone;
  const int cc = has ? c : 0;
  const int le = cc >> 6, j = (cc >> 3) & 7, k = cc & 7;
  const double* aj = s_a + (le * 8 + j) * a_stride<double>();
  const double* ak = s_a + (le * 8 + k) * a_stride<double>();
  const double on = has ? 1.0 : 0.0;
#pragma unroll 1
  for (int g = 0; g < 8; ++g) {
    const double* q = s_g + (le * 8 + g) * kKnG;
    const double wl = on * q[0], wm = on * q[1];
    const double F[9] = {q[2], q[3], q[4], q[5], q[6], q[7], q[8], q[9], q[10]};
    const double FF[9] = {q[11], q[14], q[15], q[14], q[12], q[16], q[15], q[16], q[13]};
    const double st[9] = {q[17], q[20], q[21], q[20], q[18], q[22], q[21], q[22], q[19]};
    const double a1[3] = {aj[3 * g], aj[3 * g + 1], aj[3 * g + 2]};
    const double a2[3] = {ak[3 * g], ak[3 * g + 1], ak[3 * g + 2]};
    double A1[3], A2[3], H1[3], H2[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      A1[r] = F[3 * r] * a1[0] + F[3 * r + 1] * a1[1] + F[3 * r + 2] * a1[2];
      A2[r] = F[3 * r] * a2[0] + F[3 * r + 1] * a2[1] + F[3 * r + 2] * a2[2];
      H1[r] = 0.5 * (a1[r] + A1[r]);
      H2[r] = 0.5 * (a2[r] + A2[r]);
    }
    const double d = a1[0] * a2[0] + a1[1] * a2[1] + a1[2] * a2[2];
    double gsc = 0.0;
#pragma unroll
    for (int r = 0; r < 3; ++r) gsc += a1[r] * (st[3 * r] * a2[0] + st[3 * r + 1] * a2[1] + st[3 * r + 2] * a2[2]);
    gsc *= on;
    const double wmd = wm * d;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int n = 0; n < 3; ++n) {
        KT[3 * r + n] += wl * A1[r] * A2[n] + wm * A2[r] * A1[n] + wmd * FF[3 * r + n] + (r == n ? gsc : 0.0);
        KM[3 * r + n] += wl * A1[r] * H2[n] + wm * A2[r] * H1[n] + wmd * 0.5 * (F[3 * r + n] + FF[3 * r + n]);
      }
  }
}

template <int LAYOUT>
__device__ __forceinline__ void store_raw(double* __restrict__ vals, const int32_t* s_sm, const int32_t* s_co, int out,
                                          const double v[9]) {
  const int r = out >> 8, m = out & 255;
  const int64_t sm = s_sm[r];
  const int co = s_co[r];
#pragma unroll
  for (int idx = 0; idx < 9; ++idx) vals[k_addr<LAYOUT>(sm, co, m, idx)] = v[idx];
}

template <int LAYOUT, int MODE>
__global__ void __launch_bounds__(kKnThreads)
hex8_kn_kernel(fem_mesh m, const double* __restrict__ young, double young_scalar, double lhat, double mhat,
               const double* __restrict__ u, double* __restrict__ kt, double* __restrict__ km, int eg, int rg) {
  extern __shared__ __align__(16) unsigned char smem[];
  const KnLds L = kn_lds_layout(eg, rg);
  double* s_a = reinterpret_cast<double*>(smem + L.a_off);
  double* s_g = reinterpret_cast<double*>(smem + L.g_off);
  int32_t* s_sm = reinterpret_cast<int32_t*>(smem + L.sm_off);
  int32_t* s_co = reinterpret_cast<int32_t*>(smem + L.co_off);
  const int grp = xcd_chunked(blockIdx.x, gridDim.x);
  const int e0 = m.grp_elem_ptr[grp];
  const int ne = m.grp_elem_ptr[grp + 1] - e0;
  const int r0 = m.grp_row_ptr[grp];
  const int nr = m.grp_row_ptr[grp + 1] - r0;
  const int t0 = m.grp_task_ptr[grp];
  const int nt = m.grp_task_ptr[grp + 1] - t0;
  const int tid = threadIdx.x, lane = tid & 63;
  const int le = tid >> 3, g = tid & 7;
  const bool p1 = tid < ne * 8;   // ne * 8 <= kKnThreads (checked on the host)
  for (int r = tid; r < nr; r += kKnThreads) {
    const int32_t i = m.grp_rows[r0 + r];
    s_sm[r] = m.sm_node[i];
    s_co[r] = m.co_node[i];
  }
  // ---- phase 1
  double S[6] = {0, 0, 0, 0, 0, 0};
  double det = 0.0, F[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, Ee = young_scalar;
  if (p1) {
    const int4* en = reinterpret_cast<const int4*>(m.grp_enodes + (int64_t)(e0 + le) * 8);
    const int4 na = en[0], nb = en[1];
    const int32_t nodes[8] = {na.x, na.y, na.z, na.w, nb.x, nb.y, nb.z, nb.w};
    if (young) Ee = young[m.grp_elems[e0 + le]];
    double fp[3], fm[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      fp[a] = 1.0 + gauss_sign(g, a) * kGS;
      fm[a] = 1.0 - gauss_sign(g, a) * kGS;
    }
    const int64_t n0 = nodes[0];
    const double x00 = m.coords[n0 * 3], x01 = m.coords[n0 * 3 + 1], x02 = m.coords[n0 * 3 + 2];
    double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 1; k < 8; ++k) {
      const int64_t nk = nodes[k];
      const double xr[3] = {m.coords[nk * 3] - x00, m.coords[nk * 3 + 1] - x01, m.coords[nk * 3 + 2] - x02};
      double dn[3];
      dn_factored(k, fp, fm, dn);
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) J[a * 3 + d] = fma(xr[a], dn[d], J[a * 3 + d]);
    }
    double Ji[9];
    det = inv3(J, Ji);
    double T[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};   // T[r][p] = du_r / dx_p
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      double dn[3];
      dn_factored(k, fp, fm, dn);
      double a[3];
#pragma unroll
      for (int p = 0; p < 3; ++p) a[p] = dn[0] * Ji[0 * 3 + p] + dn[1] * Ji[1 * 3 + p] + dn[2] * Ji[2 * 3 + p];
      double* dst = s_a + (le * 8 + k) * a_stride<double>() + 3 * g;
      dst[0] = a[0];
      dst[1] = a[1];
      dst[2] = a[2];
      const int64_t nk = nodes[k];
      const double uk[3] = {u[nk * 3], u[nk * 3 + 1], u[nk * 3 + 2]};
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int p = 0; p < 3; ++p) T[3 * r + p] = fma(uk[r], a[p], T[3 * r + p]);
    }
    double Eg[9];
    if (MODE == 0) {
#pragma unroll
      for (int q = 0; q < 9; ++q) F[q] = T[q] + ((q % 4 == 0) ? 1.0 : 0.0);
      // Green strain E = (F^T F - I) / 2 and the stress lambda tr(E) I + 2 mu E
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q)
          Eg[3 * p + q] = 0.5 * (F[p] * F[q] + F[3 + p] * F[3 + q] + F[6 + p] * F[6 + q] - (p == q ? 1.0 : 0.0));
    } else {
      // F stays I; the small strain sym(du/dx)
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q) Eg[3 * p + q] = 0.5 * (T[3 * p + q] + T[3 * q + p]);
    }
    const double lam = Ee * lhat, mu = Ee * mhat;
    const double tr = Eg[0] + Eg[4] + Eg[8];
    S[0] = lam * tr + 2 * mu * Eg[0];
    S[1] = lam * tr + 2 * mu * Eg[4];
    S[2] = lam * tr + 2 * mu * Eg[8];
    S[3] = 2 * mu * Eg[1];
    S[4] = 2 * mu * Eg[2];
    S[5] = 2 * mu * Eg[5];
  }
  // inclusive scan of the stress over the element's Gauss points (8 lanes)
  if (MODE == 0) {
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) {
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        const double t = __shfl_up(S[c], off, 8);
        if (g >= off) S[c] += t;
      }
    }
  }
  if (p1) {
    double* q = s_g + (le * 8 + g) * kKnG;
    q[0] = Ee * det * lhat;
    q[1] = Ee * det * mhat;
#pragma unroll
    for (int r = 0; r < 9; ++r) q[2 + r] = F[r];
    // F F^T: xx yy zz xy xz yz
    q[11] = F[0] * F[0] + F[1] * F[1] + F[2] * F[2];
    q[12] = F[3] * F[3] + F[4] * F[4] + F[5] * F[5];
    q[13] = F[6] * F[6] + F[7] * F[7] + F[8] * F[8];
    q[14] = F[0] * F[3] + F[1] * F[4] + F[2] * F[5];
    q[15] = F[0] * F[6] + F[1] * F[7] + F[2] * F[8];
    q[16] = F[3] * F[6] + F[4] * F[7] + F[5] * F[8];
#pragma unroll
    for (int c = 0; c < 6; ++c) q[17 + c] = det * S[c];
  }
  __syncthreads();
  // ---- phase 2
  for (int base = 0; base < nt; base += kKnThreads) {
    const int t = base + tid;
    const bool valid = t < nt;
    const int2 task = valid ? reinterpret_cast<const int2*>(m.tasks)[t0 + t]
                            : make_int2((int)0xFFFF0000u, kNone | (kNone << 14));
    const int outA = task.x & 0xFFFF, outB = (task.x >> 16) & 0xFFFF;
    const uint32_t w2 = (uint32_t)task.y;
    const int c1 = (int)(w2 & 0x3FFF), c2 = (int)((w2 >> 14) & 0x3FFF);
    const bool dual = (w2 >> 31) & 1u;
    const int Tn = dual ? 1 : (1 << ((w2 >> 28) & 7u));
    const bool has_b = dual && outB != 0xFFFF;
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; ++q) acc[q] = 0.0;
    kn_contrib(s_a, s_g, c1, acc, acc + 9);
    if (dual) {   // a dual task's first block is complete: store it, then reuse acc
      if (valid) {
        store_raw<LAYOUT>(kt, s_sm, s_co, outA, acc);
        if (MODE == 0) store_raw<LAYOUT>(km, s_sm, s_co, outA, acc + 9);
      }
#pragma unroll
      for (int q = 0; q < 18; ++q) acc[q] = 0.0;
    }
    kn_contrib(s_a, s_g, c2, acc, acc + 9);
    block_allsum(acc, Tn);
    if (valid && (lane & (Tn - 1)) == 0 && (!dual || has_b)) {
      store_raw<LAYOUT>(kt, s_sm, s_co, dual ? outB : outA, acc);
      if (MODE == 0) store_raw<LAYOUT>(km, s_sm, s_co, dual ? outB : outA, acc + 9);
    }
  }
}

int check_mesh(const fem_mesh* m, bool need_groups) {
  FEM_REQUIRE(m != nullptr, FEM_E_INVALID, "mesh is NULL");
  FEM_REQUIRE(m->num_node >= 0 && m->num_elem >= 0 && m->nnzb >= 0, FEM_E_INVALID, "negative size");
  FEM_REQUIRE(m->nnzb * 9 < INT32_MAX, FEM_E_LIMIT, "9*nnzb must fit int32 (got nnzb=%lld)",
              (long long)m->nnzb);
  FEM_REQUIRE(m->sm_node && m->co_node && (m->nnzb == 0 || m->re_node), FEM_E_INVALID,
              "pattern pointers are NULL");
  if (need_groups) {
    FEM_REQUIRE(m->coords && m->elem && m->grp_row_ptr && m->grp_rows && m->grp_elem_ptr &&
                    m->grp_elems && m->grp_enodes && m->grp_task_ptr &&
                    (m->max_group_tasks == 0 || m->tasks),
                FEM_E_INVALID, "mesh lacks coordinates, row groups or assembly tasks");
    FEM_REQUIRE(m->num_groups >= 0 && m->max_group_elems >= 0 && m->max_group_elems * 8 <= kAsmThreads &&
                    m->max_group_rows >= 0 && m->max_group_rows <= kAsmThreads && m->max_group_rows < 255,
                FEM_E_LIMIT, "row groups must have at most %d elements and 254 rows", kAsmThreads / 8);
    FEM_REQUIRE(((uintptr_t)m->grp_enodes & 15) == 0 && ((uintptr_t)m->tasks & 7) == 0, FEM_E_INVALID,
                "grp_enodes must be 16-byte and tasks 8-byte aligned");
  }
  return FEM_OK;
}

}  // namespace
}  // namespace fem

using namespace fem;

extern "C" int64_t fem_hex8_assemble_lds_bytes(int32_t max_group_elems, int32_t max_group_rows,
                                               int k_dtype, int want_k, int want_m) {
  if (k_dtype == FEM_DTYPE_F32) return asm_lds_layout<float>(max_group_elems, max_group_rows, want_k, want_m).total;
  return asm_lds_layout<double>(max_group_elems, max_group_rows, want_k, want_m).total;
}

extern "C" int fem_hex8_assemble(const fem_mesh* mesh, const double* young, double young_scalar,
                                 double nu, double rho, void* k_vals, int k_dtype, int k_layout,
                                 double* m_blocks, void* stream) {
  int rc = check_mesh(mesh, true);
  if (rc) return rc;
  FEM_REQUIRE(k_vals || m_blocks, FEM_E_INVALID, "nothing to assemble (k_vals and m_blocks NULL)");
  FEM_REQUIRE(k_dtype == FEM_DTYPE_F64 || k_dtype == FEM_DTYPE_F32, FEM_E_INVALID, "bad k_dtype");
  FEM_REQUIRE(k_layout == FEM_LAYOUT_BSR || k_layout == FEM_LAYOUT_CSR, FEM_E_INVALID, "bad k_layout");
  FEM_REQUIRE(nu > -1.0 && nu < 0.5, FEM_E_INVALID, "Poisson ratio %g outside (-1, 0.5)", nu);
  if (mesh->num_groups == 0) return FEM_OK;
  hipStream_t st = as_stream(stream);
  const fem_mesh& m = *mesh;
  const bool K = k_vals != nullptr, M = m_blocks != nullptr;
  if (K && k_dtype == FEM_DTYPE_F64) {
    if (k_layout == FEM_LAYOUT_BSR)
      return M ? launch_asm<double, true, true, 0>(m, young, young_scalar, nu, rho, k_vals, m_blocks, st)
               : launch_asm<double, true, false, 0>(m, young, young_scalar, nu, rho, k_vals, m_blocks, st);
    return M ? launch_asm<double, true, true, 1>(m, young, young_scalar, nu, rho, k_vals, m_blocks, st)
             : launch_asm<double, true, false, 1>(m, young, young_scalar, nu, rho, k_vals, m_blocks, st);
  }
  if (K) {
    if (k_layout == FEM_LAYOUT_BSR)
      return M ? launch_asm<float, true, true, 0>(m, young, young_scalar, nu, rho, k_vals, m_blocks, st)
               : launch_asm<float, true, false, 0>(m, young, young_scalar, nu, rho, k_vals, m_blocks, st);
    return M ? launch_asm<float, true, true, 1>(m, young, young_scalar, nu, rho, k_vals, m_blocks, st)
             : launch_asm<float, true, false, 1>(m, young, young_scalar, nu, rho, k_vals, m_blocks, st);
  }
  return launch_asm<double, false, true, 0>(m, young, young_scalar, nu, rho, nullptr, m_blocks, st);
}

extern "C" int64_t fem_hex8_kn_lds_bytes(int32_t max_group_elems, int32_t max_group_rows) {
  return kn_lds_layout(max_group_elems, max_group_rows).total;
}

static int launch_kn(const fem_mesh* mesh, const double* young, double young_scalar, double nu, const double* u,
                     double* kt_vals, double* km_vals, int k_layout, int mode, void* stream) {
  int rc = check_mesh(mesh, true);
  if (rc) return rc;
  FEM_REQUIRE(u && kt_vals && (mode == 1 || km_vals), FEM_E_INVALID, "null pointer");
  FEM_REQUIRE(k_layout == FEM_LAYOUT_BSR || k_layout == FEM_LAYOUT_CSR, FEM_E_INVALID, "bad k_layout");
  FEM_REQUIRE(nu > -1.0 && nu < 0.5, FEM_E_INVALID, "Poisson ratio %g outside (-1, 0.5)", nu);
  FEM_REQUIRE(mesh->max_group_elems * 8 <= kKnThreads, FEM_E_LIMIT,
              "large-displacement assembly needs row groups of at most %d elements", kKnThreads / 8);
  if (mesh->num_groups == 0) return FEM_OK;
  const int eg = mesh->max_group_elems, rg = mesh->max_group_rows;
  const KnLds L = kn_lds_layout(eg, rg);
  FEM_REQUIRE(L.total <= 160 * 1024, FEM_E_LIMIT, "row group needs %d B of LDS (> 160 KB)", L.total);
  const double lhat = nu / ((1 + nu) * (1 - 2 * nu));
  const double mhat = 1.0 / (2 * (1 + nu));
  hipStream_t st = as_stream(stream);
  auto kern = mode == 0 ? (k_layout == FEM_LAYOUT_BSR ? hex8_kn_kernel<0, 0> : hex8_kn_kernel<1, 0>)
                        : (k_layout == FEM_LAYOUT_BSR ? hex8_kn_kernel<0, 1> : hex8_kn_kernel<1, 1>);
  FEM_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, L.total));
  hipLaunchKernelGGL(kern, dim3((unsigned)mesh->num_groups), dim3(kKnThreads), L.total, st, *mesh, young,
                     young_scalar, lhat, mhat, u, kt_vals, km_vals, eg, rg);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_hex8_assemble_kn(const fem_mesh* mesh, const double* young, double young_scalar, double nu,
                                    const double* u, double* kt_vals, double* km_vals, int k_layout, void* stream) {
  return launch_kn(mesh, young, young_scalar, nu, u, kt_vals, km_vals, k_layout, 0, stream);
}

extern "C" int fem_hex8_assemble_ks(const fem_mesh* mesh, const double* young, double young_scalar, double nu,
                                    const double* u, double* k_vals, int k_layout, void* stream) {
  return launch_kn(mesh, young, young_scalar, nu, u, k_vals, nullptr, k_layout, 1, stream);
}

extern "C" int fem_csr_index_K(const fem_mesh* mesh, int32_t* row_k, int32_t* col_k, void* stream) {
  int rc = check_mesh(mesh, false);
  if (rc) return rc;
  FEM_REQUIRE(row_k && (mesh->nnzb == 0 || col_k), FEM_E_INVALID, "null output");
  int64_t n = mesh->nnzb > mesh->num_node ? mesh->nnzb : mesh->num_node;
  if (n == 0) {
    FEM_CHECK_HIP(hipMemsetAsync(row_k, 0, 4, as_stream(stream)));
    return FEM_OK;
  }
  csr_index_K_kernel<<<ceil_div(n, 256), 256, 0, as_stream(stream)>>>(*mesh, row_k, col_k);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_csr_index_M(const fem_mesh* mesh, int32_t* row_m, int32_t* col_m, void* stream) {
  int rc = check_mesh(mesh, false);
  if (rc) return rc;
  FEM_REQUIRE(row_m && (mesh->nnzb == 0 || col_m), FEM_E_INVALID, "null output");
  int64_t n = mesh->nnzb > mesh->num_node ? mesh->nnzb : mesh->num_node;
  if (n == 0) {
    FEM_CHECK_HIP(hipMemsetAsync(row_m, 0, 4, as_stream(stream)));
    return FEM_OK;
  }
  csr_index_M_kernel<<<ceil_div(n, 256), 256, 0, as_stream(stream)>>>(*mesh, row_m, col_m, nullptr,
                                                                     nullptr);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_csr_values_M(const fem_mesh* mesh, const double* m_blocks, double* val_m,
                                void* stream) {
  int rc = check_mesh(mesh, false);
  if (rc) return rc;
  if (mesh->nnzb == 0) return FEM_OK;
  FEM_REQUIRE(m_blocks && val_m, FEM_E_INVALID, "null pointer");
  csr_index_M_kernel<<<ceil_div(mesh->nnzb, 256), 256, 0, as_stream(stream)>>>(
      *mesh, nullptr, nullptr, m_blocks, val_m);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_spin_softening(const fem_mesh* mesh, const double* m_blocks, const double* omega3,
                                  double* k_vals, int k_layout, void* stream) {
  int rc = check_mesh(mesh, false);
  if (rc) return rc;
  FEM_REQUIRE(m_blocks && omega3 && k_vals, FEM_E_INVALID, "null pointer");
  FEM_REQUIRE(k_layout == FEM_LAYOUT_BSR || k_layout == FEM_LAYOUT_CSR, FEM_E_INVALID, "bad layout");
  if (mesh->num_node == 0) return FEM_OK;
  hipStream_t st = as_stream(stream);
  const int64_t nb = ceil_div(mesh->num_node, 256);
  if (k_layout == FEM_LAYOUT_BSR)
    spin_kernel<FEM_LAYOUT_BSR><<<nb, 256, 0, st>>>(*mesh, m_blocks, omega3[0], omega3[1], omega3[2], k_vals);
  else
    spin_kernel<FEM_LAYOUT_CSR><<<nb, 256, 0, st>>>(*mesh, m_blocks, omega3[0], omega3[1], omega3[2], k_vals);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_apply_penalty(const fem_mesh* mesh, const int32_t* bound_node,
                                 const double* bound_flag, int64_t num_bounds, void* k_vals,
                                 int k_dtype, int k_layout, double value, void* stream) {
  int rc = check_mesh(mesh, false);
  if (rc) return rc;
  if (num_bounds == 0) return FEM_OK;
  FEM_REQUIRE(num_bounds > 0 && bound_node && bound_flag && k_vals, FEM_E_INVALID, "bad arguments");
  FEM_REQUIRE(k_layout == FEM_LAYOUT_BSR || k_layout == FEM_LAYOUT_CSR, FEM_E_INVALID, "bad layout");
  hipStream_t st = as_stream(stream);
  if (k_dtype == FEM_DTYPE_F64)
    penalty_kernel<double><<<ceil_div(num_bounds, 256), 256, 0, st>>>(
        *mesh, bound_node, bound_flag, num_bounds, reinterpret_cast<double*>(k_vals), k_layout, value);
  else if (k_dtype == FEM_DTYPE_F32)
    penalty_kernel<float><<<ceil_div(num_bounds, 256), 256, 0, st>>>(
        *mesh, bound_node, bound_flag, num_bounds, reinterpret_cast<float*>(k_vals), k_layout, value);
  else
    FEM_REQUIRE(false, FEM_E_INVALID, "bad dtype");
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}
