// Load vector and stress recovery on the GPU.
//
// Both follow the same two-pass, atomic-free shape: a per-element kernel
// writes each element's per-node contributions to scratch, and a per-node
// kernel sums them over the node's incidence list (ascending element order,
// the same list the pattern builder sorted), so every result is
// deterministic.
#include "common.hpp"

namespace fem {
namespace {

// Element kernels: one lane per (element, Gauss point), 32 elements per
// work-group.  The lane computes its Gauss point's quantities into LDS; after
// a barrier lane (element, k) reduces over the 8 Gauss points for local node
// k (in Gauss order), so no lane holds an element's 8 x 3 (or 8 x 6) outputs.
constexpr int kFT = 256;
constexpr int kFE = kFT / 8;

// J at Gauss point g of the element with nodes nd, J[a][d] = dx_a/dxi_d
// (the x_0 term of sum_k x_k (x) dN_k vanishes: sum_k dN_k = 0).
__device__ __forceinline__ void jacobian_at(const fem_mesh& m, const int32_t nd[8], int g, double J[9], double R[3]) {
  const int64_t n0 = nd[0];
  const double x00 = m.coords[n0 * 3], x01 = m.coords[n0 * 3 + 1], x02 = m.coords[n0 * 3 + 2];
#pragma unroll
  for (int q = 0; q < 9; ++q) J[q] = 0.0;
  R[0] = R[1] = R[2] = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int64_t nk = nd[k];
    const double x[3] = {m.coords[nk * 3], m.coords[nk * 3 + 1], m.coords[nk * 3 + 2]};
    const double Nk = nshape(g, k);
    const double dn[3] = {dnr(g, k, 0), dnr(g, k, 1), dnr(g, k, 2)};
    const double x0[3] = {x00, x01, x02};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      R[a] = fma(Nk, x[a], R[a]);
#pragma unroll
      for (int d = 0; d < 3; ++d) J[a * 3 + d] = fma(x[a] - x0[a], dn[d], J[a * 3 + d]);
    }
  }
}

__device__ __forceinline__ void element_nodes8(const fem_mesh& m, int64_t e, int32_t nd[8]) {
  const int4* en = reinterpret_cast<const int4*>(m.elem + e * 8);
  const int4 a = en[0], b = en[1];
  nd[0] = a.x; nd[1] = a.y; nd[2] = a.z; nd[3] = a.w;
  nd[4] = b.x; nd[5] = b.y; nd[6] = b.z; nd[7] = b.w;
}

// InstallVectorF (FEM_model_csr.py:214-229) with UnitVectorF (:275-283):
// Fe[k] = sum_g rho*detJ_g * N_gk * (|w|^2 R_g - (w.R_g) w), R_g = sum_k N_gk x_k.
__global__ void __launch_bounds__(kFT)
body_force_elem(fem_mesh m, double rho, double w0, double w1, double w2, double* __restrict__ fe) {
  __shared__ double s_d[kFE][8][4];   // rho*detJ, F(R) per (element, Gauss point)
  const int le = threadIdx.x >> 3, g = threadIdx.x & 7;
  const int64_t e = (int64_t)blockIdx.x * kFE + le;
  const bool valid = e < m.num_elem;
  if (valid) {
    int32_t nd[8];
    element_nodes8(m, e, nd);
    double J[9], R[3];
    jacobian_at(m, nd, g, J, R);
    s_d[le][g][0] = rho * det3(J);
    s_d[le][g][1] = w1 * w1 * R[0] + w2 * w2 * R[0] - w0 * w1 * R[1] - w0 * w2 * R[2];
    s_d[le][g][2] = w0 * w0 * R[1] + w2 * w2 * R[1] - w1 * w2 * R[2] - w1 * w0 * R[0];
    s_d[le][g][3] = w1 * w1 * R[2] + w0 * w0 * R[2] - w2 * w0 * R[0] - w2 * w1 * R[1];
  }
  __syncthreads();
  if (valid) {
    const int k = g;   // this lane now owns local node k
    double a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const double c = s_d[le][q][0] * nshape(q, k);
      a0 = fma(c, s_d[le][q][1], a0);
      a1 = fma(c, s_d[le][q][2], a1);
      a2 = fma(c, s_d[le][q][3], a2);
    }
    fe[e * 24 + k * 3] = a0;
    fe[e * 24 + k * 3 + 1] = a1;
    fe[e * 24 + k * 3 + 2] = a2;
  }
}

// LoadF (FEM_model_csr.py:205-210) after the body force: per-node gather.
__global__ void body_force_gather(fem_mesh m, const double* __restrict__ fe,
                                  const double* __restrict__ nodal, double* __restrict__ f) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.num_node) return;
  double s0 = 0, s1 = 0, s2 = 0;
  for (int32_t t = m.ne_ptr[i]; t < m.ne_ptr[i + 1]; ++t) {
    const int64_t c = (int64_t)m.ne_list[t] * 3;   // (e*8 + j) * 3
    s0 += fe[c];
    s1 += fe[c + 1];
    s2 += fe[c + 2];
  }
  if (nodal) {
    s0 += nodal[3 * i];
    s1 += nodal[3 * i + 1];
    s2 += nodal[3 * i + 2];
  }
  f[3 * i] = s0;
  f[3 * i + 1] = s1;
  f[3 * i + 2] = s2;
}

// UnitVectorS (POST_process.py:190-209) and the extrapolation of get_NdSol
// (:172-181): per element, Gauss-point stresses, then per local node j
// sum_g shapefunc_GS(GMat[j])[g] * SVec[g] -> se[e][j][6].
__global__ void __launch_bounds__(kFT)
stress_elem(fem_mesh m, const double* __restrict__ young, double young_scalar, double lhat, double mhat,
            const double* __restrict__ u, double* __restrict__ se) {
  __shared__ double s_s[kFE][8][6];   // Gauss-point stress (xx yy zz xy yz xz)
  const int le = threadIdx.x >> 3, g = threadIdx.x & 7;
  const int64_t e = (int64_t)blockIdx.x * kFE + le;
  const bool valid = e < m.num_elem;
  if (valid) {
    int32_t nd[8];
    element_nodes8(m, e, nd);
    double J[9], R[3], Ji[9];
    jacobian_at(m, nd, g, J, R);
    inv3(J, Ji);
    const double Ee = young ? young[e] : young_scalar;
    const double lam = Ee * lhat, mu = Ee * mhat;
    double exx = 0, eyy = 0, ezz = 0, gxy = 0, gyz = 0, gxz = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const double dn[3] = {dnr(g, k, 0), dnr(g, k, 1), dnr(g, k, 2)};
      double D[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) D[a] = dn[0] * Ji[a] + dn[1] * Ji[3 + a] + dn[2] * Ji[6 + a];
      const int64_t nk = nd[k];
      const double U0 = u[nk * 3], U1 = u[nk * 3 + 1], U2 = u[nk * 3 + 2];
      exx = fma(D[0], U0, exx);
      eyy = fma(D[1], U1, eyy);
      ezz = fma(D[2], U2, ezz);
      gxy += D[1] * U0 + D[0] * U1;
      gyz += D[2] * U1 + D[1] * U2;
      gxz += D[2] * U0 + D[0] * U2;
    }
    const double tr = lam * (exx + eyy + ezz);
    s_s[le][g][0] = tr + 2 * mu * exx;
    s_s[le][g][1] = tr + 2 * mu * eyy;
    s_s[le][g][2] = tr + 2 * mu * ezz;
    s_s[le][g][3] = mu * gxy;
    s_s[le][g][4] = mu * gyz;
    s_s[le][g][5] = mu * gxz;
  }
  __syncthreads();
  if (valid) {
    const int j = g;   // this lane now owns local node j
    double out[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      // W[j][q] = shapefunc_GS(GMat[j])[q]: GMat in Gauss-point sign order,
      // weight index q taken with the shape-node sign table (as written).
      double w = 0.125;
#pragma unroll
      for (int a = 0; a < 3; ++a) w *= 1.0 + node_sign(q, a) * gauss_sign(j, a) / kGS;
#pragma unroll
      for (int c = 0; c < 6; ++c) out[c] = fma(w, s_s[le][q][c], out[c]);
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) se[e * 48 + j * 6 + c] = out[c];
  }
}

// Symmetric 3x3 eigenvalues by cyclic Jacobi rotations (FP64), descending.
__device__ void sym3_eigvals_desc(double a00, double a11, double a22, double a01, double a12,
                                  double a02, double ev[3]) {
  double A[3][3] = {{a00, a01, a02}, {a01, a11, a12}, {a02, a12, a22}};
  for (int sweep = 0; sweep < 32; ++sweep) {
    const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
    const double dg = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
    if (!(off > 1e-36 * dg) ) break;   // also exits on NaN
    for (int p = 0; p < 2; ++p) {
      for (int q = p + 1; q < 3; ++q) {
        const double apq = A[p][q];
        if (apq == 0.0) continue;
        const double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {   // rotate columns p, q
          const double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {   // rotate rows p, q
          const double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
      }
    }
  }
  double v0 = A[0][0], v1 = A[1][1], v2 = A[2][2], t;
  if (v0 < v1) { t = v0; v0 = v1; v1 = t; }
  if (v1 < v2) { t = v1; v1 = v2; v2 = t; }
  if (v0 < v1) { t = v0; v0 = v1; v1 = t; }
  ev[0] = v0; ev[1] = v1; ev[2] = v2;
}

// get_NdAve (POST_process.py:183-187) + Get_Stress host part (:79-95).
__global__ void stress_gather(fem_mesh m, const double* __restrict__ se, double* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.num_node) return;
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int32_t t = m.ne_ptr[i]; t < m.ne_ptr[i + 1]; ++t) {
    const int64_t c = (int64_t)m.ne_list[t] * 6;
#pragma unroll
    for (int q = 0; q < 6; ++q) s[q] += se[c + q];
  }
  const double ce = (double)m.co_elem[i];
#pragma unroll
  for (int q = 0; q < 6; ++q) s[q] = s[q] / ce;
  double ev[3];
  sym3_eigvals_desc(s[0], s[1], s[2], s[3], s[4], s[5], ev);
  const double vm = sqrt((ev[0] - ev[1]) * (ev[0] - ev[1]) + (ev[1] - ev[2]) * (ev[1] - ev[2]) +
                         (ev[2] - ev[1]) * (ev[2] - ev[1]));
  double* o = out + i * 10;
#pragma unroll
  for (int q = 0; q < 6; ++q) o[q] = s[q];
  o[6] = ev[0];
  o[7] = ev[1];
  o[8] = ev[2];
  o[9] = vm;
}

int check_fields_mesh(const fem_mesh* m) {
  FEM_REQUIRE(m != nullptr, FEM_E_INVALID, "mesh is NULL");
  FEM_REQUIRE(m->num_node >= 0 && m->num_elem >= 0, FEM_E_INVALID, "negative size");
  FEM_REQUIRE(m->num_node == 0 || (m->ne_ptr && m->co_elem), FEM_E_INVALID, "incidence missing");
  FEM_REQUIRE(m->num_elem == 0 || (m->coords && m->elem && m->ne_list), FEM_E_INVALID,
              "mesh arrays missing");
  return FEM_OK;
}

}  // namespace
}  // namespace fem

using namespace fem;

extern "C" int64_t fem_body_force_work_bytes(int64_t num_elem) { return 24 * 8 * num_elem + 256; }
extern "C" int64_t fem_stress_work_bytes(int64_t num_elem) { return 48 * 8 * num_elem + 256; }

extern "C" int fem_body_force(const fem_mesh* mesh, double rho, const double* omega3,
                              const double* nodal_load, double* f_out, void* work,
                              int64_t work_bytes, void* stream) {
  int rc = check_fields_mesh(mesh);
  if (rc) return rc;
  FEM_REQUIRE(omega3 != nullptr, FEM_E_INVALID, "omega3 is NULL (host pointer to 3 doubles)");
  FEM_REQUIRE(mesh->num_node == 0 || f_out, FEM_E_INVALID, "f_out is NULL");
  FEM_REQUIRE(work_bytes >= fem_body_force_work_bytes(mesh->num_elem), FEM_E_INVALID, "work too small");
  FEM_REQUIRE(((uintptr_t)mesh->elem & 15) == 0, FEM_E_INVALID, "elem must be 16-byte aligned");
  hipStream_t st = as_stream(stream);
  double* fe = reinterpret_cast<double*>(work);
  if (mesh->num_elem > 0) {
    body_force_elem<<<ceil_div(mesh->num_elem, kFE), kFT, 0, st>>>(*mesh, rho, omega3[0], omega3[1],
                                                                    omega3[2], fe);
    FEM_LAUNCH_CHECK();
  }
  if (mesh->num_node > 0) {
    body_force_gather<<<ceil_div(mesh->num_node, 256), 256, 0, st>>>(*mesh, fe, nodal_load, f_out);
    FEM_LAUNCH_CHECK();
  }
  return FEM_OK;
}

extern "C" int fem_stress_recover(const fem_mesh* mesh, const double* young, double young_scalar,
                                  double nu, const double* u, double* out, void* work,
                                  int64_t work_bytes, void* stream) {
  int rc = check_fields_mesh(mesh);
  if (rc) return rc;
  FEM_REQUIRE(mesh->num_node == 0 || (u && out), FEM_E_INVALID, "u/out NULL");
  FEM_REQUIRE(work_bytes >= fem_stress_work_bytes(mesh->num_elem), FEM_E_INVALID, "work too small");
  FEM_REQUIRE(((uintptr_t)mesh->elem & 15) == 0, FEM_E_INVALID, "elem must be 16-byte aligned");
  FEM_REQUIRE(nu > -1.0 && nu < 0.5, FEM_E_INVALID, "Poisson ratio %g outside (-1, 0.5)", nu);
  hipStream_t st = as_stream(stream);
  double* se = reinterpret_cast<double*>(work);
  const double lhat = nu / ((1 + nu) * (1 - 2 * nu));
  const double mhat = 1.0 / (2 * (1 + nu));
  if (mesh->num_elem > 0) {
    stress_elem<<<ceil_div(mesh->num_elem, kFE), kFT, 0, st>>>(*mesh, young, young_scalar, lhat, mhat,
                                                              u, se);
    FEM_LAUNCH_CHECK();
  }
  if (mesh->num_node > 0) {
    stress_gather<<<ceil_div(mesh->num_node, 256), 256, 0, st>>>(*mesh, se, out);
    FEM_LAUNCH_CHECK();
  }
  return FEM_OK;
}
