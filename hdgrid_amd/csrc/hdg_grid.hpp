// Label grid: shape, cell mapping, Morton codes.  Integer results must match the reference bit for
// bit, so the f32 arithmetic is written out operation by operation (compile with -ffp-contract=off).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace hdg {

constexpr int kMaxDims = 8;  // zorder.cuh:13 (MAX_DIMS)

// Grid state produced by the build (Super-Grid-IVF_PQ.cu:305-412).  Arrays are in ascending-span
// ("sorted") dim order; shuffle[i] is the original label column of sorted dim i.
struct GridShape {
  int l = 0;
  int dims[kMaxDims] = {0};
  uint64_t span[kMaxDims] = {0};
  uint64_t left[kMaxDims] = {0};
  int shuffle[kMaxDims] = {0};
  int max_bits = 0;
  int valid_map = 0;     // bit i: sorted dim i takes part in Morton codes and box counts
  uint64_t tot_grids = 0;
  int n_valid = 0;       // popcount(valid_map)
  int valid_dims[kMaxDims] = {0};  // sorted-dim index of the j-th valid dim
};

// calc_grid_shape restated (host).  labels [n][l] f32 host memory.  Throws on inputs the reference
// cannot represent (negative labels, zero spans).
GridShape compute_grid_shape(const float* labels, uint64_t n, int l, uint64_t max_grids,
                             uint64_t label_mask);

// f32 -> u64 as the device converts it (saturating, NaN -> 0).
__host__ __device__ inline uint64_t f2u_sat(float x) {
  if (!(x > 0.0f)) return 0;
  if (x >= 18446744073709551616.0f) return ~0ull;
  return (uint64_t)x;
}

// get_grid_id_device (Super-Grid-IVF_PQ.cu:24-26): ((u64)((x - l) * M)) / len, returned as int.
__host__ __device__ inline int grid_cell(float x, uint64_t left, uint64_t len, int M) {
  float a = x - (float)left;
  float t = a * (float)M;
  return (int)(uint32_t)(f2u_sat(t) / len);
}

// morton_encode (zorder.cu:3-15): MSB-first bit planes, sorted dims, masked dims skipped.
__host__ __device__ inline uint64_t morton_code(const int* coords, int ndims, int max_bits,
                                                int valid_map) {
  uint64_t code = 0;
  for (int b = 0; b < max_bits; ++b)
    for (int i = 0; i < ndims; ++i) {
      if (!((valid_map >> i) & 1)) continue;
      int bit = (coords[i] >> (max_bits - 1 - b)) & 1;
      code = (code << 1) | (uint64_t)bit;
    }
  return code;
}

}  // namespace hdg
