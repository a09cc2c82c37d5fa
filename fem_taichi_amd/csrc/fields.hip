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

// InstallVectorF (FEM_model_csr.py:214-229) with UnitVectorF (:275-283):
// Fe[k] = sum_g rho*detJ_g * N_gk * (|w|^2 R_g - (w.R_g) w), R_g = sum_k N_gk x_k.
__global__ void body_force_elem(fem_mesh m, double rho, double w0, double w1, double w2,
                                double* __restrict__ fe) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.num_elem) return;
  double x[8][3];
  int32_t nd[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    nd[k] = m.elem[e * 8 + k];
#pragma unroll
    for (int a = 0; a < 3; ++a) x[k][a] = m.coords[(int64_t)nd[k] * 3 + a];
  }
  double acc[8][3];
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[k][0] = acc[k][1] = acc[k][2] = 0.0;
#pragma unroll
  for (int g = 0; g < 8; ++g) {
    double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, R[3] = {0, 0, 0}, N[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      double f[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) f[a] = 1.0 + node_sign(k, a) * gauss_sign(g, a) * kGS;
      N[k] = 0.125 * f[0] * f[1] * f[2];
      double dn[3] = {0.125 * node_sign(k, 0) * f[1] * f[2], 0.125 * node_sign(k, 1) * f[0] * f[2],
                      0.125 * node_sign(k, 2) * f[0] * f[1]};
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        R[a] = fma(N[k], x[k][a], R[a]);
        const double xr = x[k][a] - x[0][a];
#pragma unroll
        for (int d = 0; d < 3; ++d) J[a * 3 + d] = fma(xr, dn[d], J[a * 3 + d]);
      }
    }
    const double F0 = w1 * w1 * R[0] + w2 * w2 * R[0] - w0 * w1 * R[1] - w0 * w2 * R[2];
    const double F1 = w0 * w0 * R[1] + w2 * w2 * R[1] - w1 * w2 * R[2] - w1 * w0 * R[0];
    const double F2 = w1 * w1 * R[2] + w0 * w0 * R[2] - w2 * w0 * R[0] - w2 * w1 * R[1];
    const double jd = rho * det3(J);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const double c = jd * N[k];
      acc[k][0] = fma(c, F0, acc[k][0]);
      acc[k][1] = fma(c, F1, acc[k][1]);
      acc[k][2] = fma(c, F2, acc[k][2]);
    }
  }
#pragma unroll
  for (int k = 0; k < 8; ++k)
#pragma unroll
    for (int a = 0; a < 3; ++a) fe[e * 24 + k * 3 + a] = acc[k][a];
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
__global__ void stress_elem(fem_mesh m, const double* __restrict__ young, double young_scalar,
                            double lhat, double mhat, const double* __restrict__ u,
                            double* __restrict__ se) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.num_elem) return;
  double x[8][3], U[8][3];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int64_t nk = m.elem[e * 8 + k];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      x[k][a] = m.coords[nk * 3 + a];
      U[k][a] = u[nk * 3 + a];
    }
  }
  const double Ee = young ? young[e] : young_scalar;
  const double lam = Ee * lhat, mu = Ee * mhat;
  double out[8][6];
#pragma unroll
  for (int j = 0; j < 8; ++j)
#pragma unroll
    for (int c = 0; c < 6; ++c) out[j][c] = 0.0;
#pragma unroll
  for (int g = 0; g < 8; ++g) {
    double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double dn[8][3];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      double f[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) f[a] = 1.0 + node_sign(k, a) * gauss_sign(g, a) * kGS;
      dn[k][0] = 0.125 * node_sign(k, 0) * f[1] * f[2];
      dn[k][1] = 0.125 * node_sign(k, 1) * f[0] * f[2];
      dn[k][2] = 0.125 * node_sign(k, 2) * f[0] * f[1];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const double xr = x[k][a] - x[0][a];
#pragma unroll
        for (int d = 0; d < 3; ++d) J[a * 3 + d] = fma(xr, dn[k][d], J[a * 3 + d]);
      }
    }
    double Ji[9];
    inv3(J, Ji);
    double exx = 0, eyy = 0, ezz = 0, gxy = 0, gyz = 0, gxz = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      double D[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) D[a] = dn[k][0] * Ji[a] + dn[k][1] * Ji[3 + a] + dn[k][2] * Ji[6 + a];
      exx = fma(D[0], U[k][0], exx);
      eyy = fma(D[1], U[k][1], eyy);
      ezz = fma(D[2], U[k][2], ezz);
      gxy += D[1] * U[k][0] + D[0] * U[k][1];
      gyz += D[2] * U[k][1] + D[1] * U[k][2];
      gxz += D[2] * U[k][0] + D[0] * U[k][2];
    }
    const double tr = lam * (exx + eyy + ezz);
    const double sv[6] = {tr + 2 * mu * exx, tr + 2 * mu * eyy, tr + 2 * mu * ezz,
                          mu * gxy, mu * gyz, mu * gxz};
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      // W[j][g] = shapefunc_GS(GMat[j])[g]: GMat in Gauss-point sign order,
      // weight index g taken with the shape-node sign table (as written).
      double w = 0.125;
#pragma unroll
      for (int a = 0; a < 3; ++a) w *= 1.0 + node_sign(g, a) * gauss_sign(j, a) / kGS;
#pragma unroll
      for (int c = 0; c < 6; ++c) out[j][c] = fma(w, sv[c], out[j][c]);
    }
  }
#pragma unroll
  for (int j = 0; j < 8; ++j)
#pragma unroll
    for (int c = 0; c < 6; ++c) se[e * 48 + j * 6 + c] = out[j][c];
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
  hipStream_t st = as_stream(stream);
  double* fe = reinterpret_cast<double*>(work);
  if (mesh->num_elem > 0) {
    body_force_elem<<<ceil_div(mesh->num_elem, 128), 128, 0, st>>>(*mesh, rho, omega3[0], omega3[1],
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
  FEM_REQUIRE(nu > -1.0 && nu < 0.5, FEM_E_INVALID, "Poisson ratio %g outside (-1, 0.5)", nu);
  hipStream_t st = as_stream(stream);
  double* se = reinterpret_cast<double*>(work);
  const double lhat = nu / ((1 + nu) * (1 - 2 * nu));
  const double mhat = 1.0 / (2 * (1 + nu));
  if (mesh->num_elem > 0) {
    stress_elem<<<ceil_div(mesh->num_elem, 64), 64, 0, st>>>(*mesh, young, young_scalar, lhat, mhat,
                                                              u, se);
    FEM_LAUNCH_CHECK();
  }
  if (mesh->num_node > 0) {
    stress_gather<<<ceil_div(mesh->num_node, 256), 256, 0, st>>>(*mesh, se, out);
    FEM_LAUNCH_CHECK();
  }
  return FEM_OK;
}
