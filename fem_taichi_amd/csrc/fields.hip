// Load vector and stress recovery on the GPU, row-owner and atomic-free.
//
// One work-group owns one row group (a spatially compact set of node rows,
// built once per mesh) and the elements touching its rows:
//   phase 1: two threads per element, four Gauss points each.  The
//            element's eight node coordinates (and displacements) are turned
//            into Walsh coefficients
//            c_S = sum_k (prod_{d in S} s_kd) v_k, S a subset of the three
//            axes and s_k the shape-node signs, so that every derivative
//            d v / d xi_d at a Gauss point is three fused multiply-adds of
//            them.  The per-Gauss-point quantities go to the element's LDS
//            slice, and after a barrier the thread pair carries them to the
//            element's nodes by a tensor-product transform (three 2-point
//            stages, one per axis); the per-node values of every element of
//            the group stay in LDS.
//   phase 2: one thread per row sums its incidences (the node -> element
//            list, ascending element order) from LDS, in a fixed order,
//            so the result is deterministic and no per-element values leave
//            the chip.
// Elements on a group's boundary are computed by every group they touch
// (about 2x for 4 x 4 x 4 node bricks), which is cheaper than a round trip
// of per-element node values through HBM.
#include "common.hpp"

namespace fem {
namespace {

constexpr int kFT = 256;   // threads per work-group: two per element

// Walsh coefficient index S: bit 0 = xi, bit 1 = eta, bit 2 = zeta.
// In-place fast transform of 8 values stored by node sign bits
// (bit d set when s_kd = +1): afterwards w[S] = sum_b prod_{d in S} (+-1) w[b].
__device__ __forceinline__ void walsh8(double w[8]) {
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int b = 0; b < 8; ++b)
      if (!((b >> d) & 1)) {
        const int b1 = b | (1 << d);
        const double lo = w[b], hi = w[b1];
        w[b] = lo + hi;
        w[b1] = hi - lo;
      }
}

// 8 * d v / d xi_d at the Gauss point with signs t0, t1, t2 (G = the
// reference's abscissa), from the Walsh coefficients c of v.
template <int D>
__device__ __forceinline__ double dxi8(const double c[8], double g0, double g1, double g2, double g01,
                                       double g02, double g12) {
  // g_a = G t_a, g_ab = G^2 t_a t_b
  if (D == 0) return fma(g12, c[7], fma(g2, c[5], fma(g1, c[3], c[1])));
  if (D == 1) return fma(g02, c[7], fma(g2, c[6], fma(g0, c[3], c[2])));
  return fma(g01, c[7], fma(g1, c[6], fma(g0, c[5], c[4])));
}

// One stage of a 2-point tensor transform along axis D over 8 values indexed
// by sign bits: (v[b0], v[b1]) -> (a v0 + b v1, b v0 + a v1), b1 = b0 | 1 << D.
template <int D>
__device__ __forceinline__ void tstage(double v[8], double a, double b) {
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (!((q >> D) & 1)) {
      const int q1 = q | (1 << D);
      const double v0 = v[q], v1 = v[q1];
      v[q] = fma(a, v0, b * v1);
      v[q1] = fma(b, v0, a * v1);
    }
}

__host__ __device__ constexpr int bitrev3(int j) { return ((j >> 2) & 1) | (j & 2) | ((j & 1) << 2); }

struct GroupRange {
  int e0, ne, r0, nr;
};

__device__ __forceinline__ GroupRange group_range(const fem_mesh& m, int g) {
  GroupRange r;
  r.e0 = m.grp_elem_ptr[g];
  r.ne = m.grp_elem_ptr[g + 1] - r.e0;
  r.r0 = m.grp_row_ptr[g];
  r.nr = m.grp_row_ptr[g + 1] - r.r0;
  return r;
}

__device__ __forceinline__ void group_element_nodes(const fem_mesh& m, int ge, int32_t nd[8]) {
  const int4* en = reinterpret_cast<const int4*>(m.grp_enodes + (int64_t)ge * 8);
  const int4 a = en[0], b = en[1];
  nd[0] = a.x; nd[1] = a.y; nd[2] = a.z; nd[3] = a.w;
  nd[4] = b.x; nd[5] = b.y; nd[6] = b.z; nd[7] = b.w;
}

// Per-element LDS slice: NC components x 8 values, padded to an odd number
// of doubles so neighbouring threads' slices start on different banks.
template <int NC>
__host__ __device__ constexpr int slice_stride() { return NC * 8 + 1; }

// LDS of a group: the element slices, then the rows' 12-int records.
template <int NC>
__host__ __device__ inline int64_t fields_lds_bytes(int eg, int rg) {
  return (int64_t)eg * slice_stride<NC>() * 8 + (int64_t)rg * 48;
}

// Row records (struct fem_mesh grp_row_rec) into LDS: issued first, so the
// row phase's metadata arrives while phase 1 runs.
__device__ __forceinline__ void load_row_record(const fem_mesh& m, int p, int4 rec[3]) {
  const int4* rp = reinterpret_cast<const int4*>(m.grp_row_rec + (int64_t)p * 12);
  rec[0] = rp[0];
  rec[1] = rp[1];
  rec[2] = rp[2];
}

// The LDS offset (in doubles) of incidence q of a row, from its record.
template <int NC, bool STRESS>
__device__ __forceinline__ int inc_offset(const fem_mesh& m, const int32_t* rec, int q) {
  const int v = q < 8 ? rec[4 + q] : m.grp_inc[rec[3] + q - 8];
  const int j = v & 7;
  return (v >> 3) * slice_stride<NC>() + (STRESS ? bitrev3(j) : node_sign_bits(j));
}

// ---------------------------------------------------------------------------
// Load vector: InstallVectorF (FEM_model_csr.py:214-229) with UnitVectorF
// (:275-283), F_k = sum_g rho detJ_g N_k(xi_g) f(R_g), R_g = sum_m N_m(xi_g)
// x_m, f(R) = |w|^2 R - (w.R) w; then LoadF (:205-210), the nodal loads.
// N_k(xi_g) = prod_d (1 + s_kd t_gd G) / 2 is a tensor product, so the
// Gauss-to-node step is three 2-point stages with a = (1 + G)/2 (same signs)
// and b = (1 - G)/2.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kFT)
body_force_rows(fem_mesh m, double rho, double w0, double w1, double w2, const double* __restrict__ nodal,
                double* __restrict__ f) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* s = reinterpret_cast<double*>(smem);
  constexpr int SS = slice_stride<3>();
  const GroupRange R = group_range(m, xcd_chunked(blockIdx.x, gridDim.x));
  int32_t* s_rec = reinterpret_cast<int32_t*>(smem + (int64_t)R.ne * SS * 8);
  const int tid = threadIdx.x;
  const bool hr = tid < R.nr;       // rows <= kFT (checked on the host)
  int4 rec[3];
  if (hr) load_row_record(m, R.r0 + tid, rec);
  const int le = tid >> 1, h = tid & 1;   // elements <= kFT / 2 (checked on the host)
  const bool he = le < R.ne;
  int32_t nd[8];
  if (he) group_element_nodes(m, R.e0 + le, nd);
  double nl0 = 0, nl1 = 0, nl2 = 0;
  if (hr) {
    reinterpret_cast<int4*>(s_rec + tid * 12)[0] = rec[0];
    reinterpret_cast<int4*>(s_rec + tid * 12)[1] = rec[1];
    reinterpret_cast<int4*>(s_rec + tid * 12)[2] = rec[2];
    if (nodal) {
      const int64_t i = rec[0].x;
      nl0 = nodal[3 * i];
      nl1 = nodal[3 * i + 1];
      nl2 = nodal[3 * i + 2];
    }
  }
  const double ww = w0 * w0 + w1 * w1 + w2 * w2;
  const double Gs = kGS;
  if (he) {
    double cx[3][8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int b = node_sign_bits(k);
      const int64_t n = nd[k];
#pragma unroll
      for (int a = 0; a < 3; ++a) cx[a][b] = m.coords[n * 3 + a];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) walsh8(cx[a]);
    double* sl = s + le * SS;
    const double g0 = h ? Gs : -Gs;   // this thread's Gauss points g = 4 h + gg
#pragma unroll
    for (int gg = 0; gg < 4; ++gg) {
      const double g1 = Gs * gauss_sign(gg, 1), g2 = Gs * gauss_sign(gg, 2);
      const double g01 = g0 * g1, g02 = g0 * g2, g12 = g1 * g2, g012 = g01 * g2;
      double J[9];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        J[a * 3 + 0] = dxi8<0>(cx[a], g0, g1, g2, g01, g02, g12);
        J[a * 3 + 1] = dxi8<1>(cx[a], g0, g1, g2, g01, g02, g12);
        J[a * 3 + 2] = dxi8<2>(cx[a], g0, g1, g2, g01, g02, g12);
      }
      const double detJ = det3(J) * (1.0 / 512.0);
      double Rg[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const double* c = cx[a];
        double v = fma(g0, c[1], c[0]);
        v = fma(g1, c[2], v);
        v = fma(g2, c[4], v);
        v = fma(g01, c[3], v);
        v = fma(g02, c[5], v);
        v = fma(g12, c[6], v);
        v = fma(g012, c[7], v);
        Rg[a] = 0.125 * v;
      }
      const double wr = w0 * Rg[0] + w1 * Rg[1] + w2 * Rg[2];
      const double sc = rho * detJ;
      const int tb = bitrev3(gg) | h;   // Gauss sign bits of g (bit 0 = xi)
      sl[0 * 8 + tb] = sc * (ww * Rg[0] - wr * w0);
      sl[8 + tb] = sc * (ww * Rg[1] - wr * w1);
      sl[16 + tb] = sc * (ww * Rg[2] - wr * w2);
    }
  }
  __syncthreads();
  // Gauss point -> node: the element's thread pair takes components h, h + 2 (of 3)
  if (he) {
    const double a = 0.5 * (1.0 + Gs), b = 0.5 * (1.0 - Gs);
    double* sl = s + le * SS;
    for (int c = h; c < 3; c += 2) {
      double v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = sl[c * 8 + q];
      tstage<0>(v, a, b);
      tstage<1>(v, a, b);
      tstage<2>(v, a, b);
#pragma unroll
      for (int q = 0; q < 8; ++q) sl[c * 8 + q] = v[q];   // by node sign bits
    }
  }
  __syncthreads();
  if (hr) {
    const int32_t* rc = s_rec + tid * 12;
    const int64_t i = rc[0];
    const int cnt = rc[2];
    double s0 = 0, s1 = 0, s2 = 0;
    for (int q = 0; q < cnt; ++q) {
      const double* sl = s + inc_offset<3, false>(m, rc, q);
      s0 += sl[0];
      s1 += sl[8];
      s2 += sl[16];
    }
    f[3 * i] = s0 + nl0;
    f[3 * i + 1] = s1 + nl1;
    f[3 * i + 2] = s2 + nl2;
  }
}

// Symmetric 3x3 eigenvalues, descending.  The trigonometric solution of the
// characteristic cubic, q + 2p cos(phi + 2k pi / 3) with q = tr A / 3, p^2 =
// |A - qI|^2 / 6 and phi = acos(det((A - qI) / p) / 2) / 3 in [0, pi/3],
// gives the root that lies alone at one end accurately (k = 0 when det >= 0,
// k = 2 otherwise: cos is flat at that root's angle), but the other two only
// to about sqrt(eps) |A| when they are close (a double root).  So only that
// root is taken from it: its eigenvector comes from the largest cross
// product of two rows of A - lambda I, A is projected onto the orthogonal
// complement, and the 2x2 projection's roots come from its entries, (m00 +
// m11)/2 +- hypot((m00 - m11)/2, m01), accurate to eps |A| however close.
__device__ void sym3_eigvals_desc(double a00, double a11, double a22, double a01, double a12, double a02,
                                  double ev[3]) {
  const double q = (a00 + a11 + a22) * (1.0 / 3.0);
  const double b00 = a00 - q, b11 = a11 - q, b22 = a22 - q;
  const double off = fma(a01, a01, fma(a02, a02, a12 * a12));
  const double p2 = fma(b00, b00, fma(b11, b11, fma(b22, b22, 2.0 * off)));
  if (!(p2 > 0.0)) {                     // A = qI (or NaN, which propagates through q)
    ev[0] = ev[1] = ev[2] = q + p2;
    return;
  }
  const double p = sqrt(p2 * (1.0 / 6.0));
  const double rp = 1.0 / p;
  const double c00 = b00 * rp, c11 = b11 * rp, c22 = b22 * rp, c01 = a01 * rp, c12 = a12 * rp, c02 = a02 * rp;
  const double detb = c00 * fma(c11, c22, -c12 * c12) - c01 * fma(c01, c22, -c12 * c02) +
                      c02 * fma(c01, c12, -c11 * c02);
  const double r = fmin(1.0, fmax(-1.0, 0.5 * detb));
  const double phi = acos(r) * (1.0 / 3.0);
  const bool top = r >= 0.0;             // the lone root is the largest (else the smallest)
  const double es = fma(2.0 * p, top ? cos(phi) : cos(phi + 2.0943951023931957), q);   // + 2 pi / 3
  // its eigenvector: the largest cross product of two rows of A - es I
  const double r0x = a00 - es, r1y = a11 - es, r2z = a22 - es;
  double vx = a01 * a12 - a02 * r1y, vy = a02 * a01 - r0x * a12, vz = r0x * r1y - a01 * a01;   // row0 x row1
  double nv = fma(vx, vx, fma(vy, vy, vz * vz));
  {
    const double x = a01 * r2z - a02 * a12, y = a02 * a02 - r0x * r2z, z = r0x * a12 - a01 * a02;  // row0 x row2
    const double n = fma(x, x, fma(y, y, z * z));
    if (n > nv) { vx = x; vy = y; vz = z; nv = n; }
  }
  {
    const double x = r1y * r2z - a12 * a12, y = a12 * a02 - a01 * r2z, z = a01 * a12 - r1y * a02;  // row1 x row2
    const double n = fma(x, x, fma(y, y, z * z));
    if (n > nv) { vx = x; vy = y; vz = z; nv = n; }
  }
  double e1, e2;
  if (!(nv > 0.0)) {                     // no usable direction: keep the cubic's other roots
    const double eo = fma(2.0 * p, top ? cos(phi + 2.0943951023931957) : cos(phi), q);
    e1 = fmax(eo, 3.0 * q - es - eo);
    e2 = fmin(eo, 3.0 * q - es - eo);
  } else {
    const double rn = 1.0 / sqrt(nv);
    vx *= rn; vy *= rn; vz *= rn;
    // orthonormal u, w spanning the complement of v
    double ux, uy, uz;
    if (fabs(vx) > fabs(vy)) {
      const double ru = 1.0 / sqrt(fma(vx, vx, vz * vz));
      ux = -vz * ru; uy = 0.0; uz = vx * ru;
    } else {
      const double ru = 1.0 / sqrt(fma(vy, vy, vz * vz));
      ux = 0.0; uy = vz * ru; uz = -vy * ru;
    }
    const double wx = vy * uz - vz * uy, wy = vz * ux - vx * uz, wz = vx * uy - vy * ux;
    const double aux = fma(a00, ux, fma(a01, uy, a02 * uz)), auy = fma(a01, ux, fma(a11, uy, a12 * uz)),
                 auz = fma(a02, ux, fma(a12, uy, a22 * uz));
    const double awx = fma(a00, wx, fma(a01, wy, a02 * wz)), awy = fma(a01, wx, fma(a11, wy, a12 * wz)),
                 awz = fma(a02, wx, fma(a12, wy, a22 * wz));
    const double m00 = fma(ux, aux, fma(uy, auy, uz * auz));
    const double m11 = fma(wx, awx, fma(wy, awy, wz * awz));
    const double m01 = fma(ux, awx, fma(uy, awy, uz * awz));
    const double h = 0.5 * (m00 + m11), d = hypot(0.5 * (m00 - m11), m01);
    e1 = h + d;
    e2 = h - d;
  }
  double v0 = top ? es : e1, v1 = top ? e1 : e2, v2 = top ? e2 : es, t;
  if (v0 < v1) { t = v0; v0 = v1; v1 = t; }
  if (v1 < v2) { t = v1; v1 = v2; v2 = t; }
  if (v0 < v1) { t = v0; v0 = v1; v1 = t; }
  ev[0] = v0;
  ev[1] = v1;
  ev[2] = v2;
}

// ---------------------------------------------------------------------------
// Stress recovery: UnitVectorS (POST_process.py:190-209) gives the stress at
// Gauss point g (RMat order), get_NdSol (:172-181) carries it to element node
// j with the weights shapefunc_GS(GMat[j])[g] = prod_d (1 + s_gd c_jd / G)/2
// (s_g the shape-node signs of index g, c_j the RMat-order corner signs of
// GMat[j]; the reference's order as written), get_NdAve (:183-187) divides
// by CoElem, and the host part of Get_Stress (:79-95) the principal stresses
// (descending) and the von Mises term sqrt((p0-p1)^2 + (p1-p2)^2 +
// (p2-p1)^2) as written.
//   grad u = H J^-1 with J = dx/dxi and H = du/dxi from the Walsh
//   coefficients; with C the cofactors of J, grad u[a][p] = H[a].C[p] / det J.
// The weights are a tensor product too: three stages with a = (1 + 1/G)/2
// (same signs) and b = (1 - 1/G)/2.  The input slot of Gauss point g is the
// sign bits of shape-node row g, the output slot of node j its RMat-order
// corner (the bit reversal of j).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kFT, 3)
stress_rows(fem_mesh m, const double* __restrict__ young, double young_scalar, double lhat, double mhat,
            const double* __restrict__ u, double* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* s = reinterpret_cast<double*>(smem);
  constexpr int SS = slice_stride<6>();
  const GroupRange R = group_range(m, xcd_chunked(blockIdx.x, gridDim.x));
  int32_t* s_rec = reinterpret_cast<int32_t*>(smem + (int64_t)R.ne * SS * 8);
  const int tid = threadIdx.x;
  const bool hr = tid < R.nr;       // rows <= kFT (checked on the host)
  int4 rec[3];
  if (hr) load_row_record(m, R.r0 + tid, rec);
  const int le = tid >> 1, h = tid & 1;   // elements <= kFT / 2 (checked on the host)
  const bool he = le < R.ne;
  int32_t nd[8];
  double Ee = young_scalar;
  if (he) {
    group_element_nodes(m, R.e0 + le, nd);
    if (young) Ee = young[m.grp_elems[R.e0 + le]];
  }
  if (hr) {
    reinterpret_cast<int4*>(s_rec + tid * 12)[0] = rec[0];
    reinterpret_cast<int4*>(s_rec + tid * 12)[1] = rec[1];
    reinterpret_cast<int4*>(s_rec + tid * 12)[2] = rec[2];
  }
  const double Gs = kGS;
  if (he) {
    double cx[3][8], cu[3][8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int b = node_sign_bits(k);
      const int64_t n = nd[k];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        cx[a][b] = m.coords[n * 3 + a];
        cu[a][b] = u[n * 3 + a];
      }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      walsh8(cx[a]);
      walsh8(cu[a]);
    }
    const double lam = Ee * lhat, mu = Ee * mhat;
    double* sl = s + le * SS;
    const double g0 = h ? Gs : -Gs;   // this thread's Gauss points g = 4 h + gg (RMat order)
#pragma unroll
    for (int gg = 0; gg < 4; ++gg) {
      const double g1 = Gs * gauss_sign(gg, 1), g2 = Gs * gauss_sign(gg, 2);
      const double g01 = g0 * g1, g02 = g0 * g2, g12 = g1 * g2;
      double J[9];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        J[a * 3 + 0] = dxi8<0>(cx[a], g0, g1, g2, g01, g02, g12);
        J[a * 3 + 1] = dxi8<1>(cx[a], g0, g1, g2, g01, g02, g12);
        J[a * 3 + 2] = dxi8<2>(cx[a], g0, g1, g2, g01, g02, g12);
      }
      // cofactors C[p][d] of J (the 8x scale of J and H cancels in H C / det)
      double C[9];
      C[0] = J[4] * J[8] - J[5] * J[7];
      C[1] = J[5] * J[6] - J[3] * J[8];
      C[2] = J[3] * J[7] - J[4] * J[6];
      C[3] = J[2] * J[7] - J[1] * J[8];
      C[4] = J[0] * J[8] - J[2] * J[6];
      C[5] = J[1] * J[6] - J[0] * J[7];
      C[6] = J[1] * J[5] - J[2] * J[4];
      C[7] = J[2] * J[3] - J[0] * J[5];
      C[8] = J[0] * J[4] - J[1] * J[3];
      const double rdet = 1.0 / (J[0] * C[0] + J[1] * C[1] + J[2] * C[2]);
      double D[9];   // D[a][p] = H[a] . C[p] = det * du_a/dx_p, H = du/dxi one row at a time
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const double h0 = dxi8<0>(cu[a], g0, g1, g2, g01, g02, g12);
        const double h1 = dxi8<1>(cu[a], g0, g1, g2, g01, g02, g12);
        const double h2 = dxi8<2>(cu[a], g0, g1, g2, g01, g02, g12);
#pragma unroll
        for (int p = 0; p < 3; ++p) D[a * 3 + p] = fma(h2, C[p * 3 + 2], fma(h1, C[p * 3 + 1], h0 * C[p * 3]));
      }
      const double l2 = lam * rdet, m2 = mu * rdet;
      const double tr = l2 * (D[0] + D[4] + D[8]);
      const int vb = node_sign_bits(4 * h + gg);   // slot: sign bits of shape-node row g
      sl[0 * 8 + vb] = fma(2.0 * m2, D[0], tr);
      sl[1 * 8 + vb] = fma(2.0 * m2, D[4], tr);
      sl[2 * 8 + vb] = fma(2.0 * m2, D[8], tr);
      sl[3 * 8 + vb] = m2 * (D[1] + D[3]);   // xy
      sl[4 * 8 + vb] = m2 * (D[5] + D[7]);   // yz
      sl[5 * 8 + vb] = m2 * (D[2] + D[6]);   // xz
    }
  }
  __syncthreads();
  // Gauss point -> node: the element's thread pair takes components h, h + 2, h + 4
  if (he) {
    const double a = 0.5 * (1.0 + 1.0 / Gs), b = 0.5 * (1.0 - 1.0 / Gs);
    double* sl = s + le * SS;
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
      const int c = 2 * cc + h;
      double v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = sl[c * 8 + q];
      tstage<0>(v, a, b);
      tstage<1>(v, a, b);
      tstage<2>(v, a, b);
#pragma unroll
      for (int q = 0; q < 8; ++q) sl[c * 8 + q] = v[q];   // by RMat-order corner bits
    }
  }
  __syncthreads();
  if (hr) {
    const int32_t* rc = s_rec + tid * 12;
    const int64_t i = rc[0];
    const double ce = (double)rc[1];
    const int cnt = rc[2];
    double sv[6] = {0, 0, 0, 0, 0, 0};
    for (int q = 0; q < cnt; ++q) {
      const double* sl = s + inc_offset<6, true>(m, rc, q);
#pragma unroll
      for (int c = 0; c < 6; ++c) sv[c] += sl[c * 8];
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) sv[c] /= ce;
    // principal stresses (descending) and the reference's von Mises term
    double ev[3];
    sym3_eigvals_desc(sv[0], sv[1], sv[2], sv[3], sv[4], sv[5], ev);
    const double vm = sqrt((ev[0] - ev[1]) * (ev[0] - ev[1]) + (ev[1] - ev[2]) * (ev[1] - ev[2]) +
                           (ev[2] - ev[1]) * (ev[2] - ev[1]));
    double* o = out + i * 10;
#pragma unroll
    for (int c = 0; c < 6; ++c) o[c] = sv[c];
    o[6] = ev[0];
    o[7] = ev[1];
    o[8] = ev[2];
    o[9] = vm;
  }
}

template <int NC>
int check_fields_mesh(const fem_mesh* m, int64_t* lds) {
  FEM_REQUIRE(m != nullptr, FEM_E_INVALID, "mesh is NULL");
  FEM_REQUIRE(m->num_node >= 0 && m->num_elem >= 0, FEM_E_INVALID, "negative size");
  FEM_REQUIRE(m->num_groups == 0 || (m->grp_row_ptr && m->grp_elem_ptr && m->grp_row_rec), FEM_E_INVALID,
              "row groups or row records (grp_row_rec) missing");
  FEM_REQUIRE(m->num_elem == 0 || (m->coords && m->grp_enodes && m->grp_elems), FEM_E_INVALID,
              "mesh arrays missing (coords, group elements)");
  FEM_REQUIRE(((uintptr_t)m->grp_enodes & 15) == 0 && ((uintptr_t)m->grp_row_rec & 15) == 0, FEM_E_INVALID,
              "grp_enodes and grp_row_rec must be 16-byte aligned");
  FEM_REQUIRE(m->max_group_elems >= 0 && m->max_group_elems <= kFT / 2, FEM_E_LIMIT,
              "row groups of %d elements: at most %d", m->max_group_elems, kFT / 2);
  FEM_REQUIRE(m->max_group_rows >= 0 && m->max_group_rows <= kFT, FEM_E_LIMIT,
              "row groups of %d rows: at most %d", m->max_group_rows, kFT);
  *lds = fields_lds_bytes<NC>(m->max_group_elems, m->max_group_rows);
  return FEM_OK;
}

}  // namespace
}  // namespace fem

using namespace fem;

extern "C" int64_t fem_body_force_work_bytes(int64_t num_elem) { return num_elem < 0 ? -1 : 0; }
extern "C" int64_t fem_stress_work_bytes(int64_t num_elem) { return num_elem < 0 ? -1 : 0; }

extern "C" int fem_body_force(const fem_mesh* mesh, double rho, const double* omega3,
                              const double* nodal_load, double* f_out, void* work,
                              int64_t work_bytes, void* stream) {
  (void)work;
  int64_t lds = 0;
  int rc = check_fields_mesh<3>(mesh, &lds);
  if (rc) return rc;
  FEM_REQUIRE(omega3 != nullptr, FEM_E_INVALID, "omega3 is NULL (host pointer to 3 doubles)");
  FEM_REQUIRE(mesh->num_node == 0 || f_out, FEM_E_INVALID, "f_out is NULL");
  FEM_REQUIRE(work_bytes >= 0, FEM_E_INVALID, "negative work_bytes");
  if (mesh->num_node == 0 || mesh->num_groups == 0) return FEM_OK;
  hipStream_t st = as_stream(stream);
  FEM_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(body_force_rows),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(body_force_rows, dim3((unsigned)mesh->num_groups), dim3(kFT), (size_t)lds, st, *mesh, rho,
                     omega3[0], omega3[1], omega3[2], nodal_load, f_out);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_stress_recover(const fem_mesh* mesh, const double* young, double young_scalar,
                                  double nu, const double* u, double* out, void* work,
                                  int64_t work_bytes, void* stream) {
  (void)work;
  int64_t lds = 0;
  int rc = check_fields_mesh<6>(mesh, &lds);
  if (rc) return rc;
  FEM_REQUIRE(mesh->num_node == 0 || (u && out), FEM_E_INVALID, "u/out NULL");
  FEM_REQUIRE(work_bytes >= 0, FEM_E_INVALID, "negative work_bytes");
  FEM_REQUIRE(nu > -1.0 && nu < 0.5, FEM_E_INVALID, "Poisson ratio %g outside (-1, 0.5)", nu);
  if (mesh->num_node == 0 || mesh->num_groups == 0) return FEM_OK;
  hipStream_t st = as_stream(stream);
  const double lhat = nu / ((1 + nu) * (1 - 2 * nu));
  const double mhat = 1.0 / (2 * (1 + nu));
  FEM_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(stress_rows),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(stress_rows, dim3((unsigned)mesh->num_groups), dim3(kFT), (size_t)lds, st, *mesh, young,
                     young_scalar, lhat, mhat, u, out);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}
