// Device top-k merge across data batches (hdg_merge.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hdg {

// dist/ids: [n_batches][nq][k] device arrays of per-data-batch results (ids batch-local); batch b covers
// rows [b * data_batch_size, min(n_data, (b + 1) * data_batch_size)).  Writes out_dist/out_ids [nq][k]:
// the k smallest per query, ascending, ties by (batch, rank), ids offset by start_offset + b *
// data_batch_size unless they are pads (>= the batch's row count).  k <= 4096.
void merge_topk(const float* dist, const uint64_t* ids, uint64_t n_batches, uint64_t data_batch_size,
                uint64_t n_data, uint64_t start_offset, uint64_t nq, uint64_t k, float* out_dist,
                uint64_t* out_ids, hipStream_t s);

}  // namespace hdg
