// hex8 stiffness / mass assembly into the node-pair pattern, row-owner and
// atomic-free.
//
// The reference (FEM_model_csr.py:150-165, 246-257; FEM_model_bsr.py:163-175)
// computes an E x 8 x 8 x 3 x 3 element intermediate and scatters it with
// atomic += after a linear slot search.  Here one work-group owns a group of
// node rows (a spatially compact set built once from the pattern):
//   phase 0: the group's row list, neighbour lists and incidence lists are
//            staged in LDS (so the rounds below touch no global memory);
//   phase 1: one thread per (element, Gauss point) of the elements touching
//            the rows computes J, stores J^-1 and w = E_e*detJ in LDS;
//   phase 2: 16 lanes own one row: lane = (local column node k, half h of
//            the Gauss points).  Round t processes the row's t-th incidence
//            (element e, local row node j) in ascending element order: the
//            lane forms the physical gradients a_gj = DNr_g[j] J_g^-1 and
//            a_gk, accumulates the 3x3 Gram block sum_g w_g a_gj (x) a_gk over
//            its four Gauss points, combines the two halves with a lane
//            shuffle and adds the block to the row's LDS accumulator at the
//            slot of column node k.  Within a round the 8 lanes of a row hit
//            8 distinct slots; rounds are ordered, so every block is a sum in
//            ascending element order -- bit-identical run to run, no atomics;
//   phase 3: every block of the group's rows is written to HBM once,
//            coalesced, as K_block = lhat*G + mhat*G^T + mhat*tr(G)*I.
// Phase 3's closed form is B_j^T D B_k with D = E*[lhat 1 1^T + 2 mhat I on
// the normal block, mhat I on the shear block] (lhat = nu/((1+nu)(1-2nu)),
// mhat = 1/(2(1+nu))): K is linear in G and E is folded into w, so the
// combination is applied once per block after the sum over elements.
#include "common.hpp"

namespace fem {
namespace {

constexpr int kAsmThreads = 256;
constexpr int kLanesPerRow = 16;
constexpr int kRowsPerPass = kAsmThreads / kLanesPerRow;

template <typename T>
struct JStride {   // per (element, Gauss point): J^-1 (9) + w (1), padded for 16-B loads
  static constexpr int value = (sizeof(T) == 8) ? 10 : 12;
};

struct AsmLds {
  int ji_off, acc_off, wm_off, accm_off, n_off, dnr_off, nodes_off, dup_off, rows_off, rowpre_off, sm_off,
      incpre_off, re_off, inc_off, slot_off, total;
};

__host__ __device__ inline int align16(int o) { return (o + 15) & ~15; }

template <typename T>
__host__ __device__ inline AsmLds asm_lds_layout(int eg, int rg, int sg, int ig, bool want_k, bool want_m) {
  AsmLds L;
  int o = 0;
  L.ji_off = o;
  o += eg * 8 * JStride<T>::value * (int)sizeof(T);
  o = align16(o);
  L.acc_off = o;
  if (want_k) o += sg * 9 * (int)sizeof(T);
  o = align16(o);
  L.wm_off = o;
  if (want_m) o += eg * 8 * (int)sizeof(double);
  o = align16(o);
  L.accm_off = o;
  if (want_m) o += sg * (int)sizeof(double);
  o = align16(o);
  L.n_off = o;
  if (want_m) o += 64 * (int)sizeof(double);
  L.dnr_off = o;
  if (want_k) o += 8 * 8 * 4 * (int)sizeof(T);   // DNr[g][k][0..2] padded to 4
  o = align16(o);
  L.nodes_off = o;
  o += eg * 8 * (int)sizeof(int32_t);
  L.dup_off = o;
  o += eg * (int)sizeof(int32_t);
  L.rows_off = o;
  o += rg * (int)sizeof(int32_t);
  L.rowpre_off = o;
  o += (rg + 1) * (int)sizeof(int32_t);
  L.sm_off = o;
  o += rg * (int)sizeof(int32_t);
  L.incpre_off = o;
  o += (rg + 1) * (int)sizeof(int32_t);
  L.re_off = o;
  o += sg * (int)sizeof(int32_t);
  L.inc_off = o;
  o += ig * (int)sizeof(int32_t);
  L.slot_off = o;
  o += ig * 8 * (int)sizeof(int16_t);
  L.total = align16(o);
  return L;
}

template <typename T, bool WANT_K, bool WANT_M, int LAYOUT>
__global__ void __launch_bounds__(kAsmThreads)
hex8_assemble_kernel(fem_mesh m, const double* __restrict__ young, double young_scalar,
                     double lhat, double mhat, double rho, T* __restrict__ k_vals,
                     double* __restrict__ m_blocks, int eg, int rg, int sg, int ig) {
  constexpr int JS = JStride<T>::value;
  extern __shared__ __align__(16) unsigned char smem[];
  const AsmLds L = asm_lds_layout<T>(eg, rg, sg, ig, WANT_K, WANT_M);
  T* s_ji = reinterpret_cast<T*>(smem + L.ji_off);
  T* s_acc = reinterpret_cast<T*>(smem + L.acc_off);
  double* s_wm = reinterpret_cast<double*>(smem + L.wm_off);
  double* s_accm = reinterpret_cast<double*>(smem + L.accm_off);
  double* s_N = reinterpret_cast<double*>(smem + L.n_off);
  T* s_dnr = reinterpret_cast<T*>(smem + L.dnr_off);
  int32_t* s_nodes = reinterpret_cast<int32_t*>(smem + L.nodes_off);
  int32_t* s_dup = reinterpret_cast<int32_t*>(smem + L.dup_off);
  int32_t* s_rows = reinterpret_cast<int32_t*>(smem + L.rows_off);
  int32_t* s_rowpre = reinterpret_cast<int32_t*>(smem + L.rowpre_off);
  int32_t* s_sm = reinterpret_cast<int32_t*>(smem + L.sm_off);
  int32_t* s_incpre = reinterpret_cast<int32_t*>(smem + L.incpre_off);
  int32_t* s_re = reinterpret_cast<int32_t*>(smem + L.re_off);
  int32_t* s_inc = reinterpret_cast<int32_t*>(smem + L.inc_off);
  int16_t* s_slot = reinterpret_cast<int16_t*>(smem + L.slot_off);

  const int grp = blockIdx.x;
  const int e0 = m.grp_elem_ptr[grp];
  const int ne = m.grp_elem_ptr[grp + 1] - e0;
  const int r0 = m.grp_row_ptr[grp];
  const int nr = m.grp_row_ptr[grp + 1] - r0;
  const int tid = threadIdx.x;

  // ---- phase 0a: per-row scalars and element node lists (independent loads)
  for (int t = tid; t < ne * 8; t += kAsmThreads) {
    const int e = m.grp_elems[e0 + (t >> 3)];
    s_nodes[t] = m.elem[(int64_t)e * 8 + (t & 7)];
  }
  for (int r = tid; r < nr; r += kAsmThreads) {
    const int32_t i = m.grp_rows[r0 + r];
    s_rows[r] = i;
    s_rowpre[r + 1] = m.co_node[i];                    // counts; scanned below
    s_sm[r] = m.sm_node[i];
    s_incpre[r + 1] = m.ne_ptr[i + 1] - m.ne_ptr[i];
  }
  if (WANT_M && tid < 64) s_N[tid] = nshape(tid >> 3, tid & 7);          // N[g*8 + k]
  if (WANT_K && tid < 192) {                                              // DNr[g][k][d]
    const int g = tid / 24, k = (tid / 3) % 8, d = tid % 3;
    s_dnr[(g * 8 + k) * 4 + d] = (T)dnr(g, k, d);
  }
  __syncthreads();
  if (tid == 0) {
    s_rowpre[0] = 0;
    s_incpre[0] = 0;
    for (int r = 0; r < nr; ++r) {
      s_rowpre[r + 1] += s_rowpre[r];
      s_incpre[r + 1] += s_incpre[r];
    }
  }
  for (int le = tid; le < ne; le += kAsmThreads) {
    int dup = 0;
    for (int a = 0; a < 8; ++a)
      for (int b = a + 1; b < 8; ++b) dup |= (s_nodes[le * 8 + a] == s_nodes[le * 8 + b]);
    s_dup[le] = dup;
  }
  __syncthreads();
  const int nslots = s_rowpre[nr];

  // ---- phase 0b: neighbour and incidence lists into LDS, zero accumulators
  for (int r = tid / kLanesPerRow; r < nr; r += kRowsPerPass) {
    const int32_t i = s_rows[r];
    const int base = s_rowpre[r], co = s_rowpre[r + 1] - base;
    for (int q = tid % kLanesPerRow; q < co; q += kLanesPerRow) s_re[base + q] = m.re_node[s_sm[r] + q];
    const int ib = m.ne_ptr[i], b2 = s_incpre[r], ni = s_incpre[r + 1] - b2;
    for (int q = tid % kLanesPerRow; q < ni; q += kLanesPerRow)
      s_inc[b2 + q] = (m.ne_local[ib + q] << 3) | (m.ne_list[ib + q] & 7);
  }
  if (WANT_K)
    for (int t = tid; t < nslots * 9; t += kAsmThreads) s_acc[t] = T(0);
  if (WANT_M)
    for (int t = tid; t < nslots; t += kAsmThreads) s_accm[t] = 0.0;
  __syncthreads();
  // slot of every (incidence, column k) in its row's sorted neighbour list --
  // the reference's linear search (FEM_model_csr.py:158-160), done once here
  {
    const int ninc = s_incpre[nr];
    for (int t = tid; t < ninc * 8; t += kAsmThreads) {
      const int inc = t >> 3, k = t & 7;
      int r = 0, hi = nr - 1;   // row of incidence inc
      while (r < hi) {
        const int mid = (r + hi + 1) >> 1;
        if (s_incpre[mid] <= inc) r = mid; else hi = mid - 1;
      }
      const int base = s_rowpre[r], co = s_rowpre[r + 1] - base;
      const int32_t c = s_nodes[(s_inc[inc] >> 3) * 8 + k];
      int lo = 0;
      hi = co - 1;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (s_re[base + mid] < c) lo = mid + 1; else hi = mid;
      }
      s_slot[t] = (int16_t)(base + lo);
    }
  }

  // ---- phase 1: J^-1 and w at each (element, Gauss point)
  for (int t = tid; t < ne * 8; t += kAsmThreads) {
    const int le = t >> 3, g = t & 7;
    const int e = m.grp_elems[e0 + le];
    double x[8][3];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int32_t nk = s_nodes[le * 8 + k];
#pragma unroll
      for (int a = 0; a < 3; ++a) x[k][a] = m.coords[(int64_t)nk * 3 + a];
    }
    double J[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) J[q] = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      double f[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) f[a] = 1.0 + node_sign(k, a) * gauss_sign(g, a) * kGS;
      const double dn[3] = {0.125 * node_sign(k, 0) * f[1] * f[2], 0.125 * node_sign(k, 1) * f[0] * f[2],
                            0.125 * node_sign(k, 2) * f[0] * f[1]};
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const double xr = x[k][a] - x[0][a];
#pragma unroll
        for (int d = 0; d < 3; ++d) J[a * 3 + d] = fma(xr, dn[d], J[a * 3 + d]);
      }
    }
    if (WANT_K) {
      double Ji[9];
      const double det = inv3(J, Ji);
      const double Ee = young ? young[e] : young_scalar;
      T* dst = s_ji + (le * 8 + g) * JS;
#pragma unroll
      for (int q = 0; q < 9; ++q) dst[q] = (T)Ji[q];
      dst[9] = (T)(Ee * det);
      if (WANT_M) s_wm[le * 8 + g] = rho * det;
    } else if (WANT_M) {
      s_wm[le * 8 + g] = rho * det3(J);
    }
  }
  __syncthreads();

  // ---- phase 2: incidence rounds, 16 lanes per row
  {
    const int lane = tid % kLanesPerRow;
    const int k = lane >> 1, h = lane & 1;
    T dk[4][3];   // DNr_g[k] for this lane's four Gauss points
    if (WANT_K) {
#pragma unroll
      for (int gg = 0; gg < 4; ++gg)
#pragma unroll
        for (int d = 0; d < 3; ++d) dk[gg][d] = s_dnr[((4 * h + gg) * 8 + k) * 4 + d];
    }
    for (int r = tid / kLanesPerRow; r < nr; r += kRowsPerPass) {
      const int ib = s_incpre[r], ie = s_incpre[r + 1];
#pragma unroll 2
      for (int t = ib; t < ie; ++t) {
        const int pk = s_inc[t];
        const int j = pk & 7, le = pk >> 3;
        const int slot = s_slot[t * 8 + k];
        T G[9];
        double ms = 0.0;
        if (WANT_K) {
#pragma unroll
          for (int q = 0; q < 9; ++q) G[q] = T(0);
#pragma unroll
          for (int gg = 0; gg < 4; ++gg) {
            const int g = 4 * h + gg;
            const T* Ji = s_ji + (le * 8 + g) * JS;
            const T* dj = s_dnr + (g * 8 + j) * 4;
            T ji[10];
#pragma unroll
            for (int q = 0; q < 10; ++q) ji[q] = Ji[q];
            const T dj0 = dj[0], dj1 = dj[1], dj2 = dj[2];
            T u[3], b[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              u[a] = ji[9] * (dj0 * ji[a] + dj1 * ji[3 + a] + dj2 * ji[6 + a]);
              b[a] = dk[gg][0] * ji[a] + dk[gg][1] * ji[3 + a] + dk[gg][2] * ji[6 + a];
            }
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
              for (int q = 0; q < 3; ++q) G[3 * p + q] = fma(u[p], b[q], G[3 * p + q]);
          }
#pragma unroll
          for (int q = 0; q < 9; ++q) G[q] += __shfl_xor(G[q], 1, 64);
        }
        if (WANT_M) {
          const double* wm = s_wm + le * 8;
#pragma unroll
          for (int gg = 0; gg < 4; ++gg) {
            const int g = 4 * h + gg;
            ms = fma(wm[g] * s_N[g * 8 + j], s_N[g * 8 + k], ms);
          }
          ms += __shfl_xor(ms, 1, 64);
        }
        if (h == 0) {
          if (!s_dup[le]) {
            if (WANT_K) {
#pragma unroll
              for (int q = 0; q < 9; ++q) s_acc[slot * 9 + q] += G[q];
            }
            if (WANT_M) s_accm[slot] += ms;
          } else {
            // a degenerate element repeats a node: add its columns one by one
            for (int kk = 0; kk < 8; ++kk) {
              if (kk == k) {
                if (WANT_K)
                  for (int q = 0; q < 9; ++q) s_acc[slot * 9 + q] += G[q];
                if (WANT_M) s_accm[slot] += ms;
              }
              __builtin_amdgcn_wave_barrier();
            }
          }
        }
      }
    }
  }
  __syncthreads();

  // ---- phase 3: write every block of the group's rows once
  if (WANT_K) {
    const T lh = (T)lhat, mh = (T)mhat;
    for (int t = tid; t < nslots * 9; t += kAsmThreads) {
      const int q9 = t / 9, e9 = t - 9 * q9;
      const int p = e9 / 3, q = e9 - 3 * p;
      int lo = 0, hi = nr - 1;   // row of slot q9: last r with s_rowpre[r] <= q9
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_rowpre[mid] <= q9) lo = mid; else hi = mid - 1;
      }
      const T* G = s_acc + q9 * 9;
      T v = lh * G[3 * p + q] + mh * G[3 * q + p];
      if (p == q) v += mh * (G[0] + G[4] + G[8]);
      const int64_t sm = s_sm[lo];
      const int64_t mloc = q9 - s_rowpre[lo];
      if (LAYOUT == FEM_LAYOUT_BSR) {
        k_vals[9 * (sm + mloc) + e9] = v;
      } else {
        const int64_t co = s_rowpre[lo + 1] - s_rowpre[lo];
        k_vals[9 * sm + 3 * p * co + 3 * mloc + q] = v;
      }
    }
  }
  if (WANT_M) {
    for (int t = tid; t < nslots; t += kAsmThreads) {
      int lo = 0, hi = nr - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_rowpre[mid] <= t) lo = mid; else hi = mid - 1;
      }
      m_blocks[(int64_t)s_sm[lo] + (t - s_rowpre[lo])] = s_accm[t];
    }
  }
}

template <typename T, bool K, bool M, int LAYOUT>
int launch_asm(const fem_mesh& m, const double* young, double ys, double nu, double rho,
               void* k_vals, double* m_blocks, hipStream_t st) {
  const int eg = m.max_group_elems, rg = m.max_group_rows, sg = m.max_group_slots, ig = m.max_group_incs;
  const AsmLds L = asm_lds_layout<T>(eg, rg, sg, ig, K, M);
  FEM_REQUIRE(L.total <= 160 * 1024, FEM_E_LIMIT,
              "row group needs %d B of LDS (> 160 KB): use smaller row groups", L.total);
  auto kern = hex8_assemble_kernel<T, K, M, LAYOUT>;
  FEM_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, L.total));
  const double lhat = nu / ((1 + nu) * (1 - 2 * nu));
  const double mhat = 1.0 / (2 * (1 + nu));
  hipLaunchKernelGGL(kern, dim3((unsigned)m.num_groups), dim3(kAsmThreads), L.total, st, m, young,
                     ys, lhat, mhat, rho, reinterpret_cast<T*>(k_vals), m_blocks, eg, rg, sg, ig);
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

int check_mesh(const fem_mesh* m, bool need_groups) {
  FEM_REQUIRE(m != nullptr, FEM_E_INVALID, "mesh is NULL");
  FEM_REQUIRE(m->num_node >= 0 && m->num_elem >= 0 && m->nnzb >= 0, FEM_E_INVALID, "negative size");
  FEM_REQUIRE(m->nnzb * 9 < INT32_MAX, FEM_E_LIMIT, "9*nnzb must fit int32 (got nnzb=%lld)",
              (long long)m->nnzb);
  FEM_REQUIRE(m->sm_node && m->co_node && (m->nnzb == 0 || m->re_node), FEM_E_INVALID,
              "pattern pointers are NULL");
  if (need_groups) {
    FEM_REQUIRE(m->coords && m->elem && m->ne_ptr && m->ne_list && m->grp_row_ptr && m->grp_rows &&
                    m->grp_elem_ptr && m->grp_elems && m->ne_local,
                FEM_E_INVALID, "mesh lacks coordinates, incidences or row groups");
    FEM_REQUIRE(m->num_groups >= 0 && m->max_group_elems >= 0 && m->max_group_rows >= 0 &&
                    m->max_group_slots >= 0 && m->max_group_incs >= 0,
                FEM_E_INVALID, "bad group sizes");
  }
  return FEM_OK;
}

}  // namespace
}  // namespace fem

using namespace fem;

extern "C" int64_t fem_hex8_assemble_lds_bytes(int32_t max_group_elems, int32_t max_group_rows,
                                               int32_t max_group_slots, int32_t max_group_incs,
                                               int k_dtype, int want_k, int want_m) {
  if (k_dtype == FEM_DTYPE_F32)
    return asm_lds_layout<float>(max_group_elems, max_group_rows, max_group_slots, max_group_incs, want_k,
                                 want_m).total;
  return asm_lds_layout<double>(max_group_elems, max_group_rows, max_group_slots, max_group_incs, want_k,
                                want_m).total;
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
