/*
 * C/OpenMP restatement of the reference's CPU stiffness / mass assembly
 * (TEST INFRASTRUCTURE ONLY: the timed CPU baseline and a second oracle).
 *
 * It follows the Taichi kernels of the reference, FP64:
 *   - FEM_CONST (FEM_model_csr.py:405-436): Gauss points in RMat order,
 *     GS = 0.57735026918962, D0/D1 blocks, per-element coordinate gather;
 *   - UnitMatrixK (:246-257) with installLocalB (:287-330): for every Gauss
 *     point J = Coord @ DNr, DN = DNr @ J^-1, and for all 64 (j, k) pairs
 *     K[j,k] += detJ * (B1_j^T D0 B1_k + B2_j^T D1 B2_k) as explicit 3x3
 *     triple products;
 *   - InstallMatrixK (:150-165): per (j, k) a linear search of the row's
 *     ReNode entries and an atomic += into ValK at
 *     9*Sm[row] + 3*p*Co[row] + 3*m + q, ColK = 3*col + q;
 *   - UnitMatrixM / InstallMatrixM (:178-192, 260-270) likewise for M.
 * The parallel loop is over elements, as Taichi's outermost loop is; the
 * element matrix is kept in a per-thread buffer instead of the reference's
 * global E x 8 x 8 x 3 x 3 field.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

static const double GSV = 0.57735026918962;
/* shape-node signs (FEM_model_csr.py:357-364) */
static const double NS[8][3] = {{1, 1, -1}, {-1, 1, -1}, {-1, -1, -1}, {1, -1, -1},
                                {1, 1, 1},  {-1, 1, 1},  {-1, -1, 1},  {1, -1, 1}};
/* Gauss point signs in RMat order (FEM_model_csr.py:408-415) */
static const double GSIGN[8][3] = {{-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1},
                                   {1, -1, -1},  {1, -1, 1},  {1, 1, -1},  {1, 1, 1}};

static void shape_derv(const double r[3], double dn[8][3]) {
  for (int k = 0; k < 8; ++k) {
    double f0 = 1 + NS[k][0] * r[0], f1 = 1 + NS[k][1] * r[1], f2 = 1 + NS[k][2] * r[2];
    dn[k][0] = 0.125 * NS[k][0] * f1 * f2;
    dn[k][1] = 0.125 * f0 * NS[k][1] * f2;
    dn[k][2] = 0.125 * f0 * f1 * NS[k][2];
  }
}

static void shape_func(const double r[3], double n[8]) {
  for (int k = 0; k < 8; ++k)
    n[k] = 0.125 * (1 + NS[k][0] * r[0]) * (1 + NS[k][1] * r[1]) * (1 + NS[k][2] * r[2]);
}

static double inv3(const double J[3][3], double Ji[3][3]) {
  double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
               J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  double r = 1.0 / det;
  Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * r;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r;
  Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * r;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r;
  Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * r;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r;
  return det;
}

/* installLocalB (FEM_model_csr.py:287-330) */
static void local_B(const double dn[3], double B1[3][3], double B2[3][3]) {
  memset(B1, 0, 9 * sizeof(double));
  memset(B2, 0, 9 * sizeof(double));
  B1[0][0] = dn[0]; B1[1][1] = dn[1]; B1[2][2] = dn[2];
  B2[0][0] = dn[1]; B2[0][1] = dn[0];
  B2[1][1] = dn[2]; B2[1][2] = dn[1];
  B2[2][0] = dn[2]; B2[2][2] = dn[0];
}

/* UnitMatrixK for one element: Ke[j][k][p][q] (FEM_model_csr.py:246-257). */
static void unit_matrix_K(const double coord[3][8], double EX, double PO, double Ke[8][8][3][3]) {
  double D0[3][3], D1[3][3];
  double a = EX * (1 - PO) / (1 + PO) / (1 - 2 * PO), b = EX * PO / (1 + PO) / (1 - 2 * PO);
  double c = EX / 2 / (1 + PO);
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q) {
      D0[p][q] = (p == q) ? a : b;
      D1[p][q] = (p == q) ? c : 0.0;
    }
  memset(Ke, 0, sizeof(double) * 8 * 8 * 9);
  for (int g = 0; g < 8; ++g) {
    double r[3] = {GSIGN[g][0] * GSV, GSIGN[g][1] * GSV, GSIGN[g][2] * GSV};
    double dnr[8][3], J[3][3] = {{0}}, Ji[3][3], DN[8][3];
    shape_derv(r, dnr);
    for (int x = 0; x < 3; ++x)
      for (int d = 0; d < 3; ++d)
        for (int k = 0; k < 8; ++k) J[x][d] += coord[x][k] * dnr[k][d];
    double det = inv3(J, Ji);
    for (int k = 0; k < 8; ++k)
      for (int x = 0; x < 3; ++x) {
        double s = 0;
        for (int d = 0; d < 3; ++d) s += dnr[k][d] * Ji[d][x];
        DN[k][x] = s;
      }
    for (int j = 0; j < 8; ++j) {
      double B11[3][3], B12[3][3];
      local_B(DN[j], B11, B12);
      for (int k = 0; k < 8; ++k) {
        double B21[3][3], B22[3][3], t1[3][3], t2[3][3];
        local_B(DN[k], B21, B22);
        /* t = D @ B2x, then Ke += det * B1x^T @ t */
        for (int p = 0; p < 3; ++p)
          for (int q = 0; q < 3; ++q) {
            double s1 = 0, s2 = 0;
            for (int m = 0; m < 3; ++m) {
              s1 += D0[p][m] * B21[m][q];
              s2 += D1[p][m] * B22[m][q];
            }
            t1[p][q] = s1;
            t2[p][q] = s2;
          }
        for (int p = 0; p < 3; ++p)
          for (int q = 0; q < 3; ++q) {
            double s = 0;
            for (int m = 0; m < 3; ++m) s += B11[m][p] * t1[m][q] + B12[m][p] * t2[m][q];
            Ke[j][k][p][q] += det * s;
          }
      }
    }
  }
}

/* InstallMatrixK (FEM_model_csr.py:150-165): CSR values (9*nnzb) and ColK.
 * coords: N x 3; elem: E x 8 node indices; young: per-element or NULL.
 * valK must be zeroed by the caller (the reference never resets it). */
void ref_assemble_K_csr(int64_t N, int64_t E, const double* coords, const int32_t* elem,
                        const double* young, double young_scalar, double nu,
                        const int32_t* sm, const int32_t* co, const int32_t* re, double* valK,
                        int32_t* colK) {
  (void)N;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t e = 0; e < E; ++e) {
    double coord[3][8];
    double Ke[8][8][3][3];
    for (int k = 0; k < 8; ++k)
      for (int x = 0; x < 3; ++x) coord[x][k] = coords[3 * (int64_t)elem[8 * e + k] + x];
    unit_matrix_K(coord, young ? young[e] : young_scalar, nu, Ke);
    for (int j = 0; j < 8; ++j) {
      int32_t row = elem[8 * e + j];
      for (int k = 0; k < 8; ++k) {
        int32_t col = elem[8 * e + k];
        for (int32_t m = 0; m < co[row]; ++m) {
          if (re[sm[row] + m] == col) {
            for (int p = 0; p < 3; ++p)
              for (int q = 0; q < 3; ++q) {
                int64_t idx = 9 * (int64_t)sm[row] + 3 * (int64_t)p * co[row] + 3 * m + q;
                if (colK) colK[idx] = 3 * col + q;
#pragma omp atomic
                valK[idx] += Ke[j][k][p][q];
              }
            break;
          }
        }
      }
    }
  }
}

/* InstallMatrixM (FEM_model_csr.py:178-192) with UnitMatrixM (:260-270):
 * valM (3*nnzb) zeroed by the caller. */
void ref_assemble_M_csr(int64_t E, const double* coords, const int32_t* elem, double rho,
                        const int32_t* sm, const int32_t* co, const int32_t* re, double* valM) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t e = 0; e < E; ++e) {
    double coord[3][8], Me[8][8];
    memset(Me, 0, sizeof(Me));
    for (int k = 0; k < 8; ++k)
      for (int x = 0; x < 3; ++x) coord[x][k] = coords[3 * (int64_t)elem[8 * e + k] + x];
    for (int g = 0; g < 8; ++g) {
      double r[3] = {GSIGN[g][0] * GSV, GSIGN[g][1] * GSV, GSIGN[g][2] * GSV};
      double dnr[8][3], nr[8], J[3][3] = {{0}}, Ji[3][3];
      shape_derv(r, dnr);
      shape_func(r, nr);
      for (int x = 0; x < 3; ++x)
        for (int d = 0; d < 3; ++d)
          for (int k = 0; k < 8; ++k) J[x][d] += coord[x][k] * dnr[k][d];
      double det = rho * inv3(J, Ji);
      for (int j = 0; j < 8; ++j)
        for (int k = 0; k < 8; ++k) Me[j][k] += det * nr[j] * nr[k];
    }
    for (int j = 0; j < 8; ++j) {
      int32_t row = elem[8 * e + j];
      for (int k = 0; k < 8; ++k) {
        int32_t col = elem[8 * e + k];
        for (int32_t m = 0; m < co[row]; ++m) {
          if (re[sm[row] + m] == col) {
            for (int p = 0; p < 3; ++p) {
              int64_t idx = 3 * (int64_t)sm[row] + (int64_t)p * co[row] + m;
#pragma omp atomic
              valM[idx] += Me[j][k];
            }
            break;
          }
        }
      }
    }
  }
}
