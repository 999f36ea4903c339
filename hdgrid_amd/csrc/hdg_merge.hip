// Top-k merge across data batches (merge_intermediate_result, merge_res.cu:18-55) on the device.
//
// Input: the per-data-batch results of one query set, [n_batches][nq][k] distances and batch-local ids, as
// the engine produces them batch after batch.  Output: per query the k smallest distances over all
// n_batches * k entries, ascending, ties by (batch, rank); each id gets its batch's global row offset
// start_offset + b * data_batch_size added in integer arithmetic, pads (ids >= the batch's row count:
// 2^32 and UINT64_MAX) pass through unchanged.  The reference adds the offset as a float, also to pads,
// which loses ids above 2^24 and wraps pads.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "hdg_common.hpp"
#include "hdg_device.hpp"
#include "hdg_merge.hpp"

namespace hdg {

namespace {

constexpr uint32_t kMergeCap = 4096;  // keys sorted in LDS (32 KiB); larger inputs are radix-selected first

// (distance bits, position b * k + j): distances are non-negative, so the bit pattern orders them
struct KeyMerge {
  const float* d;
  uint64_t nq, k, q;
  __device__ __forceinline__ uint64_t operator()(uint32_t i) const {
    const uint64_t b = i / k, j = i - b * k;
    return dev::make_key(d[(b * nq + q) * k + j], i);
  }
};

__global__ __launch_bounds__(256) void merge_topk_kernel(const float* __restrict__ dist,
                                                         const uint64_t* __restrict__ ids, uint32_t n_batches,
                                                         uint64_t data_batch_size, uint64_t n_data,
                                                         uint64_t start_offset, uint64_t nq, uint32_t k,
                                                         float* __restrict__ out_dist,
                                                         uint64_t* __restrict__ out_ids) {
  __shared__ uint64_t keys[kMergeCap];
  __shared__ __align__(16) uint32_t hist[2048];
  __shared__ uint64_t sh[8];
  const uint64_t q = blockIdx.x;
  const uint32_t n = n_batches * k;
  const KeyMerge src{dist, nq, k, q};
  uint32_t m = k;
  if (n <= kMergeCap) {
    for (uint32_t i = threadIdx.x; i < n; i += dev::kWG) keys[i] = src(i);
    __syncthreads();
    dev::wg_sort_lds(keys, n);
  } else {
    dev::wg_radix_select(src, n, k, keys, hist, sh);
    __syncthreads();
    dev::wg_sort_lds(keys, k);
  }
  __syncthreads();
  for (uint32_t r = threadIdx.x; r < m; r += dev::kWG) {
    const uint64_t key = keys[r];
    const uint32_t pos = dev::key_tie(key);
    const uint64_t b = pos / k, j = pos - b * k;
    uint64_t id = ids[(b * nq + q) * k + j];
    const uint64_t base = b * data_batch_size;
    const uint64_t rows = n_data > base ? (n_data - base < data_batch_size ? n_data - base : data_batch_size) : 0;
    if (id < rows) id += start_offset + base;
    out_dist[q * k + r] = dev::key_value(key);
    out_ids[q * k + r] = id;
  }
}

}  // namespace

void merge_topk(const float* dist, const uint64_t* ids, uint64_t n_batches, uint64_t data_batch_size,
                uint64_t n_data, uint64_t start_offset, uint64_t nq, uint64_t k, float* out_dist,
                uint64_t* out_ids, hipStream_t s) {
  if (nq == 0 || k == 0) return;
  if (!dist || !ids || !out_dist || !out_ids) fail(kInvalidArgument, "null merge pointer");
  if (n_batches == 0) fail(kInvalidArgument, "merge of zero batches");
  if (k > kMergeCap) fail(kUnsupported, "merge topk above 4096");
  if (n_batches * k >= (1ull << 32)) fail(kUnsupported, "merge input row too long");
  if (data_batch_size == 0 || (n_batches - 1) * data_batch_size >= n_data)
    fail(kInvalidArgument, "data batches do not fit n_data");
  if (nq > 0x7fffffffull) fail(kUnsupported, "too many queries for one merge");
  hipLaunchKernelGGL(merge_topk_kernel, dim3((unsigned)nq), dim3(256), 0, s, dist, ids, (uint32_t)n_batches,
                     data_batch_size, n_data, start_offset, nq, (uint32_t)k, out_dist, out_ids);
  HDG_HIP(hipGetLastError());
}

}  // namespace hdg
