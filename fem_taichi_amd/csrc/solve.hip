// BSR(3x3) SpMV / SpMM, block-Jacobi, and the fused PCG iteration kernels.
//
// SpMV: a group of 16 lanes owns one block row.  The row's 9*CoNode values
// are contiguous (BSR row-major blocks), so the 16 lanes stream them with
// unit stride; lane l handles value v = l, l+16, ... (block v/9, entry
// (v%9)/3, (v%9)%3) and the three row sums are reduced across the 16 lanes
// with shuffles.  Dot products are reduced deterministically: each
// work-group writes a partial, and the last work-group to finish (integer
// atomic ticket) sums the partials in index order.
#include "common.hpp"

namespace fem {
namespace {

constexpr int kRowLanes = 16;
constexpr int kSpmvThreads = 256;
constexpr int kRowsPerBlock = kSpmvThreads / kRowLanes;   // 16
constexpr int kVecThreads = 256;

template <typename T>
__device__ __forceinline__ T shfl_xor(T v, int mask) {
  return __shfl_xor(v, mask, 64);
}

// Sum of a double over the block (fixed tree), result valid in thread 0.
__device__ double block_sum(double v, double* red) {
  const int tid = threadIdx.x;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = tid >> 6, lane = tid & 63;
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (tid == 0) {
    const int nw = blockDim.x >> 6;
    for (int q = 0; q < nw; ++q) s += red[q];
  }
  __syncthreads();
  return s;
}

// Writes `val` as this block's partial; the last block sums all partials in
// index order into *out.  ticket must be zero on entry and is left zero.
__device__ bool finalize_sum(double val, double* partials, unsigned* ticket, double* red,
                             double* out) {
  __shared__ bool last;
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = val;
    __threadfence();
    const unsigned prev = atomicAdd(ticket, 1u);
    last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  double s = 0.0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x)
    s += __hip_atomic_load(&partials[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  s = block_sum(s, red);
  if (threadIdx.x == 0) {
    *out = s;
    *ticket = 0u;
  }
  return true;
}

template <typename T>
__device__ __forceinline__ void row_product(const fem_mesh& m, const T* __restrict__ vals,
                                            const T* __restrict__ x, int64_t i, int lane,
                                            T& a0, T& a1, T& a2) {
  a0 = a1 = a2 = T(0);
  if (i < m.num_node) {
    const int64_t sm = m.sm_node[i];
    const int nv = 9 * m.co_node[i];
    const T* __restrict__ v = vals + 9 * sm;
    const int32_t* __restrict__ cols = m.re_node + sm;
    for (int t = lane; t < nv; t += kRowLanes) {
      const int b = t / 9, r = t - 9 * b;
      const int p = r / 3, q = r - 3 * p;
      const T prod = v[t] * x[3 * (int64_t)cols[b] + q];
      a0 += (p == 0) ? prod : T(0);
      a1 += (p == 1) ? prod : T(0);
      a2 += (p == 2) ? prod : T(0);
    }
  }
#pragma unroll
  for (int o = kRowLanes / 2; o > 0; o >>= 1) {
    a0 += shfl_xor(a0, o);
    a1 += shfl_xor(a1, o);
    a2 += shfl_xor(a2, o);
  }
}

// y = A x; optionally partial x.y for the PCG step (state[1] = pq,
// state[2] = alpha = state[0] / pq).
template <typename T, bool DOT>
__global__ void __launch_bounds__(kSpmvThreads)
spmv_kernel(fem_mesh m, const T* __restrict__ vals, const T* __restrict__ x, T* __restrict__ y,
            double* partials, unsigned* ticket, double* state) {
  __shared__ double red[kSpmvThreads / 64];
  const int lane = threadIdx.x & (kRowLanes - 1);
  const int64_t i = (int64_t)blockIdx.x * kRowsPerBlock + (threadIdx.x / kRowLanes);
  T a0, a1, a2;
  row_product(m, vals, x, i, lane, a0, a1, a2);
  double d = 0.0;
  if (i < m.num_node && lane == 0) {
    y[3 * i] = a0;
    y[3 * i + 1] = a1;
    y[3 * i + 2] = a2;
    if (DOT) d = (double)a0 * (double)x[3 * i] + (double)a1 * (double)x[3 * i + 1] +
                 (double)a2 * (double)x[3 * i + 2];
  }
  if (DOT) {
    d = block_sum(d, red);
    double pq;
    if (finalize_sum(d, partials, ticket, red, &state[1])) {
      if (threadIdx.x == 0) {
        pq = state[1];
        state[2] = state[0] / pq;
      }
    }
  }
}

// SpMM: Y = A X, X/Y (3N, nrhs) row-major.  Thread per (block row, column).
__global__ void spmm_kernel(fem_mesh m, const double* __restrict__ vals, const double* __restrict__ X,
                            double* __restrict__ Y, int nrhs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t / nrhs;
  const int c = (int)(t - i * nrhs);
  if (i >= m.num_node) return;
  const int64_t sm = m.sm_node[i];
  const int co = m.co_node[i];
  double a0 = 0, a1 = 0, a2 = 0;
  for (int b = 0; b < co; ++b) {
    const double* v = vals + 9 * (sm + b);
    const int64_t col = m.re_node[sm + b];
    const double x0 = X[(3 * col) * nrhs + c], x1 = X[(3 * col + 1) * nrhs + c],
                 x2 = X[(3 * col + 2) * nrhs + c];
    a0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
    a1 += v[3] * x0 + v[4] * x1 + v[5] * x2;
    a2 += v[6] * x0 + v[7] * x1 + v[8] * x2;
  }
  Y[(3 * i) * nrhs + c] = a0;
  Y[(3 * i + 1) * nrhs + c] = a1;
  Y[(3 * i + 2) * nrhs + c] = a2;
}

// Mass product with one scalar per node-pair block on each DOF diagonal.
__global__ void mass_spmm_kernel(fem_mesh m, const double* __restrict__ mb, const double* __restrict__ X,
                                 double* __restrict__ Y, int nrhs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t / nrhs;
  const int c = (int)(t - i * nrhs);
  if (i >= m.num_node) return;
  const int64_t sm = m.sm_node[i];
  const int co = m.co_node[i];
  double a0 = 0, a1 = 0, a2 = 0;
  for (int b = 0; b < co; ++b) {
    const double w = mb[sm + b];
    const int64_t col = m.re_node[sm + b];
    a0 += w * X[(3 * col) * nrhs + c];
    a1 += w * X[(3 * col + 1) * nrhs + c];
    a2 += w * X[(3 * col + 2) * nrhs + c];
  }
  Y[(3 * i) * nrhs + c] = a0;
  Y[(3 * i + 1) * nrhs + c] = a1;
  Y[(3 * i + 2) * nrhs + c] = a2;
}

// Inverse of each node's diagonal 3x3 block (FP64 arithmetic).
template <typename T>
__global__ void bj_setup_kernel(fem_mesh m, const T* __restrict__ vals, T* __restrict__ dinv,
                                int32_t* flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.num_node) return;
  const int64_t sm = m.sm_node[i];
  const int co = m.co_node[i];
  int k = -1;
  for (int q = 0; q < co; ++q)
    if (m.re_node[sm + q] == i) { k = q; break; }
  double B[9], Bi[9];
  bool ok = k >= 0;
  if (ok) {
    for (int q = 0; q < 9; ++q) B[q] = (double)vals[9 * (sm + k) + q];
    const double det = inv3(B, Bi);
    ok = (det != 0.0) && isfinite(det);
    for (int q = 0; q < 9; ++q) ok = ok && isfinite(Bi[q]);
  }
  if (!ok) {
    for (int q = 0; q < 9; ++q) Bi[q] = (q % 4 == 0) ? 1.0 : 0.0;
    if (flag) atomicMax(flag, 1);
  }
  for (int q = 0; q < 9; ++q) dinv[9 * i + q] = (T)Bi[q];
}

// r = b - q; z = Dinv r; p = z; partial sums of r.z, r.r, b.b.  q may alias z
// (each thread reads its own entries before writing them).
template <typename T>
__global__ void __launch_bounds__(kVecThreads)
pcg_start_kernel(int64_t n, const T* __restrict__ b, const T* q, const T* __restrict__ dinv,
                 T* r, T* z, T* p, double* partials,
                 unsigned* tickets, double* state) {
  __shared__ double red[kVecThreads / 64];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rz = 0, rr = 0, bb = 0;
  if (i < n) {
    T ri[3], zi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) ri[a] = b[3 * i + a] - q[3 * i + a];
    const T* D = dinv + 9 * i;
#pragma unroll
    for (int a = 0; a < 3; ++a) zi[a] = D[3 * a] * ri[0] + D[3 * a + 1] * ri[1] + D[3 * a + 2] * ri[2];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      r[3 * i + a] = ri[a];
      z[3 * i + a] = zi[a];
      p[3 * i + a] = zi[a];
      rz += (double)ri[a] * (double)zi[a];
      rr += (double)ri[a] * (double)ri[a];
      const double bv = (double)b[3 * i + a];
      bb += bv * bv;
    }
  }
  const int G = gridDim.x;
  rz = block_sum(rz, red);
  finalize_sum(rz, partials, tickets, red, &state[0]);
  rr = block_sum(rr, red);
  finalize_sum(rr, partials + G, tickets + 1, red, &state[4]);
  bb = block_sum(bb, red);
  finalize_sum(bb, partials + 2 * G, tickets + 2, red, &state[5]);
}

// x += a p; r -= a q; z = Dinv r; partial r.z, r.r -> beta.
template <typename T>
__global__ void __launch_bounds__(kVecThreads)
pcg_update_kernel(int64_t n, T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                  const T* __restrict__ p, const T* __restrict__ q, const T* __restrict__ dinv,
                  double* partials, unsigned* tickets, double* state) {
  __shared__ double red[kVecThreads / 64];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double alpha = state[2];
  const T a = (T)alpha;
  double rz = 0, rr = 0;
  if (i < n) {
    T ri[3], zi[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      x[3 * i + c] += a * p[3 * i + c];
      ri[c] = r[3 * i + c] - a * q[3 * i + c];
      r[3 * i + c] = ri[c];
    }
    const T* D = dinv + 9 * i;
#pragma unroll
    for (int c = 0; c < 3; ++c) zi[c] = D[3 * c] * ri[0] + D[3 * c + 1] * ri[1] + D[3 * c + 2] * ri[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      z[3 * i + c] = zi[c];
      rz += (double)ri[c] * (double)zi[c];
      rr += (double)ri[c] * (double)ri[c];
    }
  }
  const int G = gridDim.x;
  rz = block_sum(rz, red);
  double rz_old = state[0];
  if (finalize_sum(rz, partials, tickets, red, &state[6])) {
    if (threadIdx.x == 0) {
      state[3] = state[6] / rz_old;
      state[0] = state[6];
    }
  }
  rr = block_sum(rr, red);
  finalize_sum(rr, partials + G, tickets + 1, red, &state[4]);
}

template <typename T>
__global__ void pcg_direction_kernel(int64_t n3, const T* __restrict__ z, T* __restrict__ p,
                                     const double* state) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n3) return;
  const T beta = (T)state[3];
  p[t] = z[t] + beta * p[t];
}

int check_solve_mesh(const fem_mesh* m) {
  FEM_REQUIRE(m != nullptr, FEM_E_INVALID, "mesh is NULL");
  FEM_REQUIRE(m->num_node >= 0 && m->nnzb >= 0, FEM_E_INVALID, "negative size");
  FEM_REQUIRE(m->num_node == 0 || (m->sm_node && m->co_node), FEM_E_INVALID, "pattern missing");
  FEM_REQUIRE(m->nnzb == 0 || m->re_node, FEM_E_INVALID, "re_node missing");
  return FEM_OK;
}

struct PcgWork {
  double* partials;
  unsigned* tickets;
};

int64_t vec_blocks(int64_t N) { return ceil_div(N > 0 ? N : 1, kVecThreads); }
int64_t spmv_blocks(int64_t N) { return ceil_div(N > 0 ? N : 1, kRowsPerBlock); }

PcgWork pcg_work(void* work, int64_t N) {
  int64_t nb = spmv_blocks(N) > vec_blocks(N) ? spmv_blocks(N) : vec_blocks(N);
  PcgWork w;
  w.tickets = reinterpret_cast<unsigned*>(work);
  w.partials = reinterpret_cast<double*>(reinterpret_cast<char*>(work) + 256);
  (void)nb;
  return w;
}

}  // namespace
}  // namespace fem

using namespace fem;

extern "C" int64_t fem_pcg_work_bytes(int64_t num_node) {
  int64_t nb = spmv_blocks(num_node) > vec_blocks(num_node) ? spmv_blocks(num_node) : vec_blocks(num_node);
  return 256 + 3 * 8 * nb + 256;
}

extern "C" int fem_bsr_spmv(const fem_mesh* mesh, const void* vals, int dtype, const void* x, void* y,
                            void* stream) {
  int rc = check_solve_mesh(mesh);
  if (rc) return rc;
  if (mesh->num_node == 0) return FEM_OK;
  FEM_REQUIRE(vals && x && y, FEM_E_INVALID, "null pointer");
  hipStream_t st = as_stream(stream);
  const int64_t nb = spmv_blocks(mesh->num_node);
  if (dtype == FEM_DTYPE_F64)
    spmv_kernel<double, false><<<nb, kSpmvThreads, 0, st>>>(
        *mesh, (const double*)vals, (const double*)x, (double*)y, nullptr, nullptr, nullptr);
  else if (dtype == FEM_DTYPE_F32)
    spmv_kernel<float, false><<<nb, kSpmvThreads, 0, st>>>(
        *mesh, (const float*)vals, (const float*)x, (float*)y, nullptr, nullptr, nullptr);
  else
    FEM_REQUIRE(false, FEM_E_INVALID, "bad dtype");
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_bsr_spmm(const fem_mesh* mesh, const double* vals, const double* X, double* Y,
                            int64_t nrhs, void* stream) {
  int rc = check_solve_mesh(mesh);
  if (rc) return rc;
  FEM_REQUIRE(nrhs > 0 && nrhs <= 1024, FEM_E_INVALID, "nrhs must be in [1, 1024]");
  if (mesh->num_node == 0) return FEM_OK;
  FEM_REQUIRE(vals && X && Y, FEM_E_INVALID, "null pointer");
  const int64_t T = mesh->num_node * nrhs;
  spmm_kernel<<<ceil_div(T, 256), 256, 0, as_stream(stream)>>>(*mesh, vals, X, Y, (int)nrhs);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_mass_spmm(const fem_mesh* mesh, const double* m_blocks, const double* X, double* Y,
                             int64_t nrhs, void* stream) {
  int rc = check_solve_mesh(mesh);
  if (rc) return rc;
  FEM_REQUIRE(nrhs > 0 && nrhs <= 1024, FEM_E_INVALID, "nrhs must be in [1, 1024]");
  if (mesh->num_node == 0) return FEM_OK;
  FEM_REQUIRE(m_blocks && X && Y, FEM_E_INVALID, "null pointer");
  const int64_t T = mesh->num_node * nrhs;
  mass_spmm_kernel<<<ceil_div(T, 256), 256, 0, as_stream(stream)>>>(*mesh, m_blocks, X, Y, (int)nrhs);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_bj_setup(const fem_mesh* mesh, const void* vals, int dtype, void* dinv, int32_t* flag,
                            void* stream) {
  int rc = check_solve_mesh(mesh);
  if (rc) return rc;
  if (mesh->num_node == 0) return FEM_OK;
  FEM_REQUIRE(vals && dinv, FEM_E_INVALID, "null pointer");
  hipStream_t st = as_stream(stream);
  const int64_t nb = ceil_div(mesh->num_node, 256);
  if (dtype == FEM_DTYPE_F64)
    bj_setup_kernel<double><<<nb, 256, 0, st>>>(*mesh, (const double*)vals, (double*)dinv, flag);
  else if (dtype == FEM_DTYPE_F32)
    bj_setup_kernel<float><<<nb, 256, 0, st>>>(*mesh, (const float*)vals, (float*)dinv, flag);
  else
    FEM_REQUIRE(false, FEM_E_INVALID, "bad dtype");
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

template <typename T>
static int pcg_start_t(const fem_mesh* mesh, const T* vals, const T* dinv, const T* b, T* x, T* r, T* z,
                       T* p, double* state, void* work, hipStream_t st) {
  const int64_t N = mesh->num_node;
  PcgWork w = pcg_work(work, N);
  FEM_CHECK_HIP(hipMemsetAsync(w.tickets, 0, 256, st));
  // q := A x (into z as temporary), then r = b - q etc.
  spmv_kernel<T, false><<<spmv_blocks(N), kSpmvThreads, 0, st>>>(*mesh, vals, x, z, nullptr, nullptr,
                                                                  nullptr);
  pcg_start_kernel<T><<<vec_blocks(N), kVecThreads, 0, st>>>(N, b, z, dinv, r, z, p, w.partials,
                                                            w.tickets, state);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

template <typename T>
static int pcg_iterate_t(const fem_mesh* mesh, const T* vals, const T* dinv, T* x, T* r, T* z, T* p,
                         T* q, double* state, void* work, hipStream_t st) {
  const int64_t N = mesh->num_node;
  PcgWork w = pcg_work(work, N);
  spmv_kernel<T, true><<<spmv_blocks(N), kSpmvThreads, 0, st>>>(*mesh, vals, p, q, w.partials,
                                                                 w.tickets, state);
  pcg_update_kernel<T><<<vec_blocks(N), kVecThreads, 0, st>>>(N, x, r, z, p, q, dinv, w.partials,
                                                             w.tickets + 4, state);
  pcg_direction_kernel<T><<<ceil_div(3 * N, 256), 256, 0, st>>>(3 * N, z, p, state);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

extern "C" int fem_pcg_start(const fem_mesh* mesh, const void* vals, const void* dinv, int dtype,
                             const void* b, void* x, void* r, void* z, void* p, double* state,
                             void* work, int64_t work_bytes, void* stream) {
  int rc = check_solve_mesh(mesh);
  if (rc) return rc;
  if (mesh->num_node == 0) return FEM_OK;
  FEM_REQUIRE(vals && dinv && b && x && r && z && p && state && work, FEM_E_INVALID, "null pointer");
  FEM_REQUIRE(work_bytes >= fem_pcg_work_bytes(mesh->num_node), FEM_E_INVALID, "work too small");
  hipStream_t st = as_stream(stream);
  if (dtype == FEM_DTYPE_F64)
    return pcg_start_t<double>(mesh, (const double*)vals, (const double*)dinv, (const double*)b,
                               (double*)x, (double*)r, (double*)z, (double*)p, state, work, st);
  if (dtype == FEM_DTYPE_F32)
    return pcg_start_t<float>(mesh, (const float*)vals, (const float*)dinv, (const float*)b, (float*)x,
                              (float*)r, (float*)z, (float*)p, state, work, st);
  FEM_REQUIRE(false, FEM_E_INVALID, "bad dtype");
}

extern "C" int fem_pcg_iterate(const fem_mesh* mesh, const void* vals, const void* dinv, int dtype,
                               void* x, void* r, void* z, void* p, void* q, double* state, void* work,
                               int64_t work_bytes, void* stream) {
  int rc = check_solve_mesh(mesh);
  if (rc) return rc;
  if (mesh->num_node == 0) return FEM_OK;
  FEM_REQUIRE(vals && dinv && x && r && z && p && q && state && work, FEM_E_INVALID, "null pointer");
  FEM_REQUIRE(work_bytes >= fem_pcg_work_bytes(mesh->num_node), FEM_E_INVALID, "work too small");
  hipStream_t st = as_stream(stream);
  if (dtype == FEM_DTYPE_F64)
    return pcg_iterate_t<double>(mesh, (const double*)vals, (const double*)dinv, (double*)x, (double*)r,
                                 (double*)z, (double*)p, (double*)q, state, work, st);
  if (dtype == FEM_DTYPE_F32)
    return pcg_iterate_t<float>(mesh, (const float*)vals, (const float*)dinv, (float*)x, (float*)r,
                                (float*)z, (float*)p, (float*)q, state, work, st);
  FEM_REQUIRE(false, FEM_E_INVALID, "bad dtype");
}
