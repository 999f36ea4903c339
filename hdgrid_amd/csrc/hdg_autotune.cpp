// Launch-parameter autotuner: coordinate descent over the query knobs that do not change results.
//
// The knobs: ADC_PER, ADC_WGS, QUERY_CHUNK, SCORE_BLOCK, ADC_SPLIT and GRAPH never change a result;
// SPARSE_MAX_TESTS does not under PRECISION 0, but under PRECISION 1 it decides which queries refine
// straight from their in-box lists (instead of the FP16-LUT Morton route), so it is tuned only under
// PRECISION 0.  Each knob is tried over a fixed ladder of values with the others held at the best setting so far; a
// candidate is timed as the median of `reps` full queries (HIP events on the caller's stream, one
// untimed query first, which also captures the graphs of GRAPH 1) and accepted only when it is faster
// AND its ids and distances equal those of the starting configuration bit for bit.  The final setting
// stays in the index's configuration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "hdg_config.hpp"
#include "hdg_index.hpp"

namespace hdg {

namespace {

struct Knob {
  const char* key;
  std::vector<uint64_t> values;
};

}  // namespace

std::string Index::autotune(const float* queries, uint64_t nq, uint64_t d, const float* qlabels, uint64_t qcols,
                            float* out_dist, uint64_t* out_ids, uint32_t reps, hipStream_t s) {
  HDG_HIP(hipSetDevice(device_));
  if (!st_.built) fail(kBadState, "autotune before build");
  if (nq == 0) fail(kInvalidArgument, "autotune needs queries");
  reps = std::max<uint32_t>(1, std::min<uint32_t>(reps, 64));
  const uint64_t k = cfg_.topk;
  hipEvent_t e0, e1;
  HDG_HIP(hipEventCreate(&e0));
  HDG_HIP(hipEventCreate(&e1));
  std::vector<float> ref_d(nq * k), got_d(nq * k);
  std::vector<uint64_t> ref_i(nq * k), got_i(nq * k);
  auto fetch = [&](std::vector<float>& dd, std::vector<uint64_t>& ii) {
    HDG_HIP(hipMemcpyAsync(dd.data(), out_dist, nq * k * 4, hipMemcpyDeviceToHost, s));
    HDG_HIP(hipMemcpyAsync(ii.data(), out_ids, nq * k * 8, hipMemcpyDeviceToHost, s));
    HDG_HIP(hipStreamSynchronize(s));
  };
  auto time_current = [&]() -> float {
    query(queries, nq, d, qlabels, qcols, out_dist, out_ids, s);
    std::vector<float> t(reps);
    for (uint32_t r = 0; r < reps; ++r) {
      HDG_HIP(hipEventRecord(e0, s));
      query(queries, nq, d, qlabels, qcols, out_dist, out_ids, s);
      HDG_HIP(hipEventRecord(e1, s));
      HDG_HIP(hipEventSynchronize(e1));
      HDG_HIP(hipEventElapsedTime(&t[r], e0, e1));
    }
    std::nth_element(t.begin(), t.begin() + reps / 2, t.end());
    return t[reps / 2];
  };
  auto set = [&](const char* key, uint64_t v) {
    if (!set_config_value(cfg_, key, std::to_string(v))) fail(kConfigError, std::string("unknown key ") + key);
  };
  auto get = [&](const char* key) {
    uint64_t v = 0;
    get_config_value(cfg_, key, v);
    return v;
  };

  std::string report;
  char line[160];
  float best;
  try {
    best = time_current();
    fetch(ref_d, ref_i);
    std::snprintf(line, sizeof line, "start %.4f ms\n", best);
    report += line;
    const uint64_t chunk_cap = std::max<uint64_t>(4096, nq);
    std::vector<Knob> knobs = {
        {"ADC_PER", {8, 16, 32, 64}},
        {"ADC_WGS", {512, 1024, 2048, 4096, 8192}},
        {"SPARSE_MAX_TESTS", {16384, 32768, 65536, 131072, 262144, 524288}},
        {"QUERY_CHUNK", {}},
        {"SCORE_BLOCK", {512, 1024, 2048, 4096, 16384}},
        {"ADC_SPLIT", {0, 16384, 65536, 262144}},
        {"GRAPH", {0, 1}},
    };
    for (uint64_t c = 2048; c <= 65536 && c <= chunk_cap * 2; c *= 2) knobs[3].values.push_back(c);
    if (cfg_.precision != 0) knobs[2].values.clear();  // result-changing under PRECISION 1
    for (const Knob& kb : knobs) {
      const uint64_t keep = get(kb.key);
      uint64_t best_v = keep;
      for (uint64_t v : kb.values) {
        if (v == keep) continue;
        set(kb.key, v);
        const float t = time_current();
        fetch(got_d, got_i);
        const bool same = std::memcmp(got_d.data(), ref_d.data(), nq * k * 4) == 0 &&
                          std::memcmp(got_i.data(), ref_i.data(), nq * k * 8) == 0;
        std::snprintf(line, sizeof line, "%s=%llu %.4f ms%s\n", kb.key, (unsigned long long)v, t,
                      same ? "" : " (results differ: rejected)");
        report += line;
        if (same && t < best) {
          best = t;
          best_v = v;
        }
      }
      set(kb.key, best_v);
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::snprintf(line, sizeof line, "best %.4f ms:", best);
  report += line;
  for (const char* key : {"ADC_PER", "ADC_WGS", "SPARSE_MAX_TESTS", "QUERY_CHUNK", "SCORE_BLOCK", "ADC_SPLIT", "GRAPH"}) {
    std::snprintf(line, sizeof line, " %s=%llu", key, (unsigned long long)get(key));
    report += line;
  }
  report += "\n";
  return report;
}

}  // namespace hdg
