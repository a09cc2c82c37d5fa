// Shared helpers for the fem355 HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <cstring>

#include "../../include/fem355.h"

namespace fem {

// Thread-local last-error message returned by fem_last_error().
void set_error(const char* fmt, ...);

#define FEM_CHECK_HIP(expr)                                                       \
  do {                                                                           \
    hipError_t _e = (expr);                                                      \
    if (_e != hipSuccess) {                                                      \
      fem::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #expr,                \
                     hipGetErrorString(_e));                                     \
      return FEM_E_HIP;                                                          \
    }                                                                            \
  } while (0)

#define FEM_REQUIRE(cond, code, ...)                                             \
  do {                                                                           \
    if (!(cond)) {                                                               \
      fem::set_error(__VA_ARGS__);                                               \
      return (code);                                                             \
    }                                                                            \
  } while (0)

#define FEM_LAUNCH_CHECK()                                                       \
  do {                                                                           \
    hipError_t _e = hipGetLastError();                                           \
    if (_e != hipSuccess) {                                                      \
      fem::set_error("%s:%d kernel launch -> %s", __FILE__, __LINE__,            \
                     hipGetErrorString(_e));                                     \
      return FEM_E_HIP;                                                          \
    }                                                                            \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Work-groups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8).  Give each XCD one contiguous run of the (spatially ordered) row
// groups so that the coordinates neighbouring groups share stay in its L2.
__device__ __forceinline__ int xcd_chunked(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, x = b & 7, i = b >> 3;
  return x * q + min(x, r) + i;
}

// Gauss abscissa exactly as written in the reference (FEM_model_csr.py:93).
constexpr double kGS = 0.57735026918962;

// Shape-function node signs (FEM_model_csr.py:357-364), packed as bit d set
// when the sign along axis d is +1: k0(+,+,-) k1(-,+,-) k2(-,-,-) k3(+,-,-)
// k4(+,+,+) k5(-,+,+) k6(-,-,+) k7(+,-,+).
__host__ __device__ constexpr int node_sign_bits(int k) {
  return (k == 0) ? 3 : (k == 1) ? 2 : (k == 2) ? 0 : (k == 3) ? 1
       : (k == 4) ? 7 : (k == 5) ? 6 : (k == 6) ? 4 : 5;
}
__host__ __device__ constexpr double node_sign(int k, int d) {
  return ((node_sign_bits(k) >> d) & 1) ? 1.0 : -1.0;
}
// Gauss point order (FEM_model_csr.py:408-415): g = 4*bx + 2*by + bz, bit
// set means +GS.  Sign along axis d of Gauss point g.
__host__ __device__ constexpr double gauss_sign(int g, int d) {
  return ((g >> (2 - d)) & 1) ? 1.0 : -1.0;
}

// dN_k/dxi_d at Gauss point g (FEM_model_csr.py:368-402).
__host__ __device__ inline double dnr(int g, int k, int d) {
  double f[3];
  for (int a = 0; a < 3; ++a) f[a] = 1.0 + node_sign(k, a) * gauss_sign(g, a) * kGS;
  double v = 0.125 * node_sign(k, d);
  for (int a = 0; a < 3; ++a)
    if (a != d) v *= f[a];
  return v;
}
// N_k at Gauss point g (FEM_model_csr.py:355-365).
__host__ __device__ inline double nshape(int g, int k) {
  double v = 0.125;
  for (int a = 0; a < 3; ++a) v *= 1.0 + node_sign(k, a) * gauss_sign(g, a) * kGS;
  return v;
}

// 3x3 inverse and determinant by cofactors, J row-major J[a][d] = dx_a/dxi_d.
template <typename T>
__device__ __forceinline__ T inv3(const T J[9], T Ji[9]) {
  T c00 = J[4] * J[8] - J[5] * J[7];
  T c01 = J[5] * J[6] - J[3] * J[8];
  T c02 = J[3] * J[7] - J[4] * J[6];
  T det = J[0] * c00 + J[1] * c01 + J[2] * c02;
  T r = T(1) / det;
  Ji[0] = c00 * r;
  Ji[1] = (J[2] * J[7] - J[1] * J[8]) * r;
  Ji[2] = (J[1] * J[5] - J[2] * J[4]) * r;
  Ji[3] = c01 * r;
  Ji[4] = (J[0] * J[8] - J[2] * J[6]) * r;
  Ji[5] = (J[2] * J[3] - J[0] * J[5]) * r;
  Ji[6] = c02 * r;
  Ji[7] = (J[1] * J[6] - J[0] * J[7]) * r;
  Ji[8] = (J[0] * J[4] - J[1] * J[3]) * r;
  return det;
}

template <typename T>
__device__ __forceinline__ T det3(const T J[9]) {
  return J[0] * (J[4] * J[8] - J[5] * J[7]) + J[1] * (J[5] * J[6] - J[3] * J[8]) +
         J[2] * (J[3] * J[7] - J[4] * J[6]);
}

}  // namespace fem
