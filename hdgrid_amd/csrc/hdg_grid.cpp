// Grid shape: Super-Grid-IVF_PQ.cu:305-412 (calc_grid_shape) with the min/max of
// preprocessing.cu:175-204 (calculate_batch_min_max).
#include "hdg_grid.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <utility>

#include "hdg_common.hpp"

namespace hdg {

GridShape compute_grid_shape(const float* labels, uint64_t n, int l, uint64_t max_grids,
                             uint64_t label_mask_cfg) {
  if (l < 1 || l > kMaxDims) fail(kInvalidArgument, "label dims must be in [1, 8]");
  if (n == 0) fail(kInvalidArgument, "empty dataset");
  std::vector<float> lo(l, FLT_MAX), hi(l, -FLT_MAX);
  for (uint64_t p = 0; p < n; ++p)
    for (int i = 0; i < l; ++i) {
      float v = labels[p * l + i];
      if (v < lo[i]) lo[i] = v;
      if (v > hi[i]) hi[i] = v;
    }
  GridShape g;
  g.l = l;
  const uint8_t active = (uint8_t)(~(uint8_t)label_mask_cfg);
  int cnt = 0;
  for (int i = 0; i < l; ++i) cnt += (active >> i) & 1;
  if (cnt == 0) fail(kInvalidArgument, "label_mask masks every label dim");
  // (span, original dim) sorted ascending; span = ceilf(max) - floorf(min) in f32
  std::vector<std::pair<uint64_t, int>> order(l);
  for (int i = 0; i < l; ++i) {
    if (lo[i] < 0.0f) fail(kInvalidArgument, "negative label values are not supported");
    float s = std::ceil(hi[i]) - std::floor(lo[i]);
    order[i] = {f2u_host(s), i};
  }
  std::sort(order.begin(), order.end());
  int main_grids = (int)std::floor(std::sqrt((double)max_grids));
  if (l == 1) main_grids = (int)std::min<uint64_t>(max_grids, order[0].first);
  int remain = 1, remain_cnt = cnt - 1, last_valid = l - 1;
  if (l > 1) remain = (int)(max_grids / (uint64_t)main_grids);
  uint8_t morton_mask = 0xFF;
  for (int i = 0; i < l; ++i) {
    const int id = order[i].second;
    const uint64_t len = order[i].first;
    if ((active >> id) & 1) {
      if (remain_cnt == 0) {
        last_valid = i;
        continue;
      }
      const uint64_t root =
          (uint64_t)std::floor(std::pow((double)main_grids, 1.0 / (double)(float)(cnt - 1)));
      int len_i = (int)std::min<uint64_t>((uint64_t)remain, std::min<uint64_t>(len, root));
      if (remain_cnt == 1) len_i = (int)std::min<uint64_t>((uint64_t)remain, len);
      remain_cnt--;
      if (len_i <= 0) fail(kInvalidArgument, "label column with zero span");
      g.dims[i] = len_i;
      remain /= len_i;
    } else {
      morton_mask &= (uint8_t)~(1u << i);
      g.dims[i] = 1;
    }
  }
  if (order[last_valid].first == 0 || main_grids <= 0)
    fail(kInvalidArgument, "label column with zero span");
  g.dims[last_valid] = main_grids;
  g.tot_grids = 1;
  for (int i = 0; i < l; ++i) g.tot_grids *= (uint64_t)g.dims[i];
  for (int i = 0; i < l; ++i) {
    const int id = order[i].second;
    g.span[i] = order[i].first;
    g.left[i] = f2u_host(std::floor(lo[id]));
    g.shuffle[i] = id;
  }
  int maxd = 0;
  for (int i = 0; i < l; ++i) maxd = std::max(maxd, g.dims[i]);
  g.max_bits = (int)std::ceil(std::log2((double)(maxd + 1)));
  g.valid_map = (int)(uint8_t)(morton_mask & (uint8_t)((1ull << l) - 1));
  g.n_valid = 0;
  for (int i = 0; i < l; ++i)
    if ((g.valid_map >> i) & 1) g.valid_dims[g.n_valid++] = i;
  if (g.max_bits * g.n_valid > 64) fail(kUnsupported, "Morton code wider than 64 bits");
  return g;
}

}  // namespace hdg
