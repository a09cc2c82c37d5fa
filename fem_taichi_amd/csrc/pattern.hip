// Symbolic node-pair pattern on the GPU: the arrays READ_physics.py:42-61
// derives with SciPy (Elem_Node = COO->CSC, Node_Node = Elem_Node^T Elem_Node,
// sort_indices).  Integer work only: counts use integer atomics (whose
// result is order independent) and every list is sorted afterwards, so the
// output is bit-identical run to run and to SciPy's.
#include "common.hpp"

namespace fem {
namespace {

constexpr int kScanBlock = 1024;
constexpr int kMaxNbr = 128;    // distinct neighbour nodes of one node (per-thread buffer)

// Flags any element-node entry outside [0, N) (checked before any scatter).
__global__ void check_range(const int32_t* __restrict__ elem, int64_t T, int64_t N,
                            int32_t* __restrict__ err) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  int32_t v = elem[t];
  if (v < 0 || v >= N) atomicMax(err, 2);
}

// Count (e, j) incidences per node (all of them, for ne_list) and distinct
// elements per node (CoElem: COO->CSC sums duplicate (e, node) entries).
__global__ void count_incidence(const int32_t* __restrict__ elem, int64_t E,
                                int32_t* __restrict__ n_inc, int32_t* __restrict__ co_elem) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= E * 8) return;
  int64_t e = t >> 3;
  int j = (int)(t & 7);
  int32_t node = elem[t];
  atomicAdd(&n_inc[node], 1);
  bool first = true;
  for (int jj = 0; jj < j; ++jj) first &= (elem[e * 8 + jj] != node);
  if (first) atomicAdd(&co_elem[node], 1);
}

// Block-level exclusive scan of int32 -> writes block totals.
__global__ void scan_blocks(const int32_t* __restrict__ in, int64_t n, int32_t* __restrict__ out,
                            int32_t* __restrict__ block_sums) {
  __shared__ int32_t s[kScanBlock];
  int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
  int32_t v = (i < n) ? in[i] : 0;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    int32_t t = (threadIdx.x >= (unsigned)off) ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) out[i] = s[threadIdx.x] - v;
  if (threadIdx.x == kScanBlock - 1) block_sums[blockIdx.x] = s[threadIdx.x];
}

// Sequential exclusive scan of the (few thousand) block totals by one block.
__global__ void scan_totals(int32_t* __restrict__ block_sums, int64_t nb, int32_t* __restrict__ total) {
  __shared__ int32_t s[kScanBlock];
  __shared__ int32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < nb; base += kScanBlock) {
    int64_t i = base + threadIdx.x;
    int32_t v = (i < nb) ? block_sums[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
      int32_t t = (threadIdx.x >= (unsigned)off) ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) block_sums[i] = carry + s[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == kScanBlock - 1) carry += s[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void scan_add(int32_t* __restrict__ out, int64_t n, const int32_t* __restrict__ block_sums,
                         const int32_t* __restrict__ total) {
  int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
  if (i < n) out[i] += block_sums[blockIdx.x];
  if (i == 0) out[n] = *total;
}

__global__ void fill_incidence(const int32_t* __restrict__ elem, int64_t E,
                               const int32_t* __restrict__ ne_ptr, int32_t* __restrict__ cursor,
                               int32_t* __restrict__ ne_list) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= E * 8) return;
  int32_t node = elem[t];
  int32_t pos = atomicAdd(&cursor[node], 1);
  ne_list[ne_ptr[node] + pos] = (int32_t)t;   // e*8 + j
}

// Sort each node's incidence list (short) so its order is (e, j) ascending.
__global__ void sort_incidence(const int32_t* __restrict__ ne_ptr, int64_t N, int32_t* __restrict__ ne_list) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int32_t b = ne_ptr[i], e = ne_ptr[i + 1];
  for (int32_t a = b + 1; a < e; ++a) {
    int32_t v = ne_list[a];
    int32_t c = a - 1;
    while (c >= b && ne_list[c] > v) {
      ne_list[c + 1] = ne_list[c];
      --c;
    }
    ne_list[c + 1] = v;
  }
}

// Sorted unique neighbours of node i (union of its elements' nodes).
// mode 0: count into co_node; mode 1: write to re_node[sm_node[i]...].
template <int MODE>
__global__ void neighbours(const int32_t* __restrict__ elem, int64_t N,
                           const int32_t* __restrict__ ne_ptr, const int32_t* __restrict__ ne_list,
                           int32_t* __restrict__ co_node, const int32_t* __restrict__ sm_node,
                           int32_t* __restrict__ re_node, int32_t* __restrict__ err) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int32_t b = ne_ptr[i], e = ne_ptr[i + 1];
  int32_t buf[kMaxNbr];
  int n = 0;
  bool overflow = false;
  for (int32_t t = b; t < e && !overflow; ++t) {
    int32_t el = ne_list[t] >> 3;
    for (int k = 0; k < 8; ++k) {
      int32_t c = elem[(int64_t)el * 8 + k];
      // insert c into sorted buf[0..n) if absent
      int pos = n;
      while (pos > 0 && buf[pos - 1] > c) --pos;
      if (pos > 0 && buf[pos - 1] == c) continue;
      if (n == kMaxNbr) {
        overflow = true;
        break;
      }
      for (int q = n; q > pos; --q) buf[q] = buf[q - 1];
      buf[pos] = c;
      ++n;
    }
  }
  if (overflow) {
    if (MODE == 0) {
      atomicMax(err, 1);
      co_node[i] = 0;
    }
    return;
  }
  if (MODE == 0) {
    co_node[i] = n;
  } else {
    int32_t o = sm_node[i];
    for (int q = 0; q < n; ++q) re_node[o + q] = buf[q];
  }
}

int exclusive_scan(const int32_t* in, int64_t n, int32_t* out, int32_t* block_sums, int32_t* total,
                   hipStream_t st) {
  int64_t nb = ceil_div(n > 0 ? n : 1, kScanBlock);
  scan_blocks<<<nb, kScanBlock, 0, st>>>(in, n, out, block_sums);
  scan_totals<<<1, kScanBlock, 0, st>>>(block_sums, nb, total);
  scan_add<<<nb, kScanBlock, 0, st>>>(out, n, block_sums, total);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}

}  // namespace
}  // namespace fem

using namespace fem;

extern "C" int64_t fem_pattern_work_bytes(int64_t num_elem, int64_t num_node) {
  int64_t nb = ceil_div(num_node > 0 ? num_node : 1, kScanBlock);
  // n_inc (N) + cursor (N) + block sums + total + err, int32 each, 256-B aligned chunks
  return 4 * (2 * num_node + nb + 64) + 1024;
}

extern "C" int fem_pattern_count(const int32_t* elem, int64_t num_elem, int64_t num_node,
                                 int32_t* co_elem, int32_t* ne_ptr, int32_t* ne_list,
                                 int32_t* co_node, int32_t* sm_node, void* work,
                                 int64_t work_bytes, void* stream) {
  FEM_REQUIRE(num_elem >= 0 && num_node >= 0, FEM_E_INVALID, "negative size");
  FEM_REQUIRE(num_elem * 8 < INT32_MAX, FEM_E_LIMIT, "8*E must fit int32");
  FEM_REQUIRE(co_elem && ne_ptr && co_node && sm_node && work, FEM_E_INVALID, "null pointer");
  FEM_REQUIRE(num_elem == 0 || (elem && ne_list), FEM_E_INVALID, "null pointer");
  FEM_REQUIRE(work_bytes >= fem_pattern_work_bytes(num_elem, num_node), FEM_E_INVALID,
              "work too small: %lld < %lld", (long long)work_bytes,
              (long long)fem_pattern_work_bytes(num_elem, num_node));
  hipStream_t st = as_stream(stream);
  int64_t nb = ceil_div(num_node > 0 ? num_node : 1, kScanBlock);
  int32_t* n_inc = reinterpret_cast<int32_t*>(work);
  int32_t* cursor = n_inc + num_node;
  int32_t* bsums = cursor + num_node;
  int32_t* total = bsums + nb;
  int32_t* err = total + 16;
  FEM_CHECK_HIP(hipMemsetAsync(work, 0, 4 * (2 * num_node + nb + 64), st));
  FEM_CHECK_HIP(hipMemsetAsync(co_elem, 0, 4 * num_node, st));
  if (num_elem > 0) {
    int64_t T = num_elem * 8;
    int32_t herr0 = 0;
    check_range<<<ceil_div(T, 256), 256, 0, st>>>(elem, T, num_node, err);
    FEM_LAUNCH_CHECK();
    FEM_CHECK_HIP(hipMemcpyAsync(&herr0, err, 4, hipMemcpyDeviceToHost, st));
    FEM_CHECK_HIP(hipStreamSynchronize(st));
    FEM_REQUIRE(herr0 == 0, FEM_E_INVALID, "element connectivity holds a node index outside [0, %lld)",
                (long long)num_node);
    count_incidence<<<ceil_div(T, 256), 256, 0, st>>>(elem, num_elem, n_inc, co_elem);
    FEM_LAUNCH_CHECK();
  }
  int rc = exclusive_scan(n_inc, num_node, ne_ptr, bsums, total, st);
  if (rc) return rc;
  if (num_elem > 0) {
    int64_t T = num_elem * 8;
    fill_incidence<<<ceil_div(T, 256), 256, 0, st>>>(elem, num_elem, ne_ptr, cursor, ne_list);
    FEM_LAUNCH_CHECK();
  }
  if (num_node > 0) {
    sort_incidence<<<ceil_div(num_node, 256), 256, 0, st>>>(ne_ptr, num_node, ne_list);
    neighbours<0><<<ceil_div(num_node, 128), 128, 0, st>>>(elem, num_node, ne_ptr, ne_list, co_node,
                                                           nullptr, nullptr, err);
    FEM_LAUNCH_CHECK();
  }
  rc = exclusive_scan(co_node, num_node, sm_node, bsums, total, st);
  if (rc) return rc;
  int32_t herr = 0;
  FEM_CHECK_HIP(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, st));
  FEM_CHECK_HIP(hipStreamSynchronize(st));
  FEM_REQUIRE(herr == 0, FEM_E_LIMIT,
              "a node has more than %d distinct neighbour nodes (pattern builder limit)", kMaxNbr);
  return FEM_OK;
}

extern "C" int fem_pattern_fill(const int32_t* elem, int64_t num_node, const int32_t* ne_ptr,
                                const int32_t* ne_list, const int32_t* sm_node, int32_t* re_node,
                                void* stream) {
  FEM_REQUIRE(num_node >= 0, FEM_E_INVALID, "negative size");
  if (num_node == 0) return FEM_OK;
  FEM_REQUIRE(ne_ptr && sm_node, FEM_E_INVALID, "null pointer");
  hipStream_t st = as_stream(stream);
  neighbours<1><<<ceil_div(num_node, 128), 128, 0, st>>>(elem, num_node, ne_ptr, ne_list, nullptr,
                                                         sm_node, re_node, nullptr);
  FEM_LAUNCH_CHECK();
  return FEM_OK;
}
