// C ABI (include/hdgrid.h) over hdg::Index.  Every entry point converts exceptions into status codes
// and a thread-local message; nothing here exits the process.
#include "../../include/hdgrid.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <string>

#include "hdg_fast.hpp"
#include "hdg_index.hpp"
#include "hdg_merge.hpp"

struct hdg_index {
  hdg::Index* impl;
};

namespace {

thread_local std::string g_last_error;

template <class F>
hdg_status guarded(F&& f) {
  try {
    f();
    g_last_error.clear();
    return HDG_OK;
  } catch (const hdg::Error& e) {
    g_last_error = e.what();
    return e.status;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return HDG_ENOMEM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return HDG_EINVAL;
  }
}

// index-factory.cuh:32,84 accept only "super-grid"; the README spells it "Super-Grid" -- accept both.
void check_algorithm(const char* algorithm) {
  if (!algorithm) hdg::fail(HDG_ECONFIG, "null algorithm name");
  std::string a(algorithm);
  std::transform(a.begin(), a.end(), a.begin(), [](unsigned char c) { return (char)std::tolower(c); });
  if (a != "super-grid") hdg::fail(HDG_ECONFIG, "Unknown algorithm: " + std::string(algorithm));
}

// IS_SPLIT = 1 selects the Morton split mode (zorder.cu:474-654) with N_SPLIT pieces per range; the range table
// of a query holds SUB_LISTS * N_SPLIT pieces.
void check_supported(const hdg::Config& cfg) {
  if (cfg.is_split && (cfg.n_split == 0 || cfg.n_split * cfg.sub_lists > 8000))
    hdg::fail(HDG_EUNSUPPORTED, "IS_SPLIT = 1 needs N_SPLIT >= 1 and SUB_LISTS * N_SPLIT <= 8000");
}

void check_device(int device) {
  int n = 0;
  HDG_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) hdg::fail(HDG_EINVAL, "device ordinal out of range");
}

// split_task restated (search-engine-auto-batch.cu:430-489).  The reference's bisection never ends
// when no query batch of >= 125 fits; this one reports that instead.
void split_task(const double* coeff, uint64_t n_data, uint64_t n_queries, uint64_t topk, uint64_t mem_bound,
                uint64_t available, uint64_t data_width, uint64_t& qb, uint64_t& db) {
  const uint64_t upper = std::min(available / 5, mem_bound);
  uint64_t r_d = n_data, l_d = topk, l_q = 1, r_q;
  if (n_data < topk) hdg::fail(HDG_EINVAL, "n_data < topk");
  while (true) {
    const uint64_t mid_d = (r_d + l_d + 1) >> 1;
    r_q = n_queries;
    l_q = 1;
    while (l_q < r_q) {
      const uint64_t mid_q = (l_q + r_q + 1) >> 1;
      const double v = coeff[0] * (double)mid_d + coeff[1] * (double)mid_q +
                       coeff[2] * (double)mid_d * (double)mid_q + coeff[3] + 1;
      const uint64_t value = v <= 0 ? 0 : (v >= 1.8446744073709552e19 ? ~0ull : (uint64_t)v);
      if (value > upper) r_q = mid_q - 1; else l_q = mid_q;
    }
    if (l_q < 125) {
      if (mid_d == 0 || mid_d - 1 < l_d) hdg::fail(HDG_EINVAL, "no data batch admits a query batch of 125");
      r_d = mid_d - 1;
    } else {
      l_d = mid_d;
    }
    if (l_d == r_d && l_q >= 125) break;
  }
  qb = 125;
  db = (uint64_t)hdg::find_max_factor(l_d, n_data);
  qb = (uint64_t)hdg::find_max_factor(qb, n_queries);
  if (db * 4 <= n_data) db *= data_width;
}

// gauss (math_utils.h:6-41) restated: Gauss-Jordan elimination with partial pivoting on the augmented
// n x (n+1) system, pivots below 1e-30 skipped; returns 0 (unique solution left in column n), 1 (rank
// deficient, consistent) or 2 (inconsistent).
int gauss(double a[][5], int n) {
  const double eps = 1e-30;
  int r = 0;
  for (int c = 0; c < n; ++c) {
    int t = r;
    for (int i = r; i < n; ++i)
      if (std::fabs(a[i][c]) > std::fabs(a[t][c])) t = i;
    if (std::fabs(a[t][c]) < eps) continue;
    for (int i = c; i <= n; ++i) std::swap(a[t][i], a[r][i]);
    for (int i = n; i >= c; --i) a[r][i] /= a[r][c];
    for (int i = r + 1; i < n; ++i)
      if (std::fabs(a[i][c]) > eps)
        for (int j = n; j >= c; --j) a[i][j] -= a[i][c] * a[r][j];
    ++r;
  }
  if (r < n) {
    for (int i = r; i < n; ++i)
      if (std::fabs(a[i][n]) > eps) return 2;
    return 1;
  }
  for (int i = n - 1; i >= 0; --i)
    for (int j = i + 1; j < n; ++j) a[i][n] -= a[j][n] * a[i][j];
  return 0;
}

}  // namespace

extern "C" {

const char* hdg_last_error(void) { return g_last_error.c_str(); }

hdg_status hdg_create(const char* algorithm, const char* conf_path, const char* filter_conf_path, int device,
                      hdg_index** out) {
  return guarded([&] {
    if (!out || !conf_path) hdg::fail(HDG_EINVAL, "null argument");
    *out = nullptr;
    check_algorithm(algorithm);
    hdg::Config cfg;
    hdg::parse_config_file(conf_path, cfg);
    check_supported(cfg);
    check_device(device);
    std::string fpath = filter_conf_path && *filter_conf_path ? std::string(filter_conf_path)
                                                               : cfg.path + "filter.conf";
    hdg::FilterConfig fc;
    hdg::parse_filter_file(fpath, fc);
    *out = new hdg_index{new hdg::Index(cfg, fc, device)};
  });
}

hdg_status hdg_create_from_text(const char* algorithm, const char* conf_text, const char* filter_text, int device,
                                hdg_index** out) {
  return guarded([&] {
    if (!out || !conf_text || !filter_text) hdg::fail(HDG_EINVAL, "null argument");
    *out = nullptr;
    check_algorithm(algorithm);
    hdg::Config cfg;
    hdg::parse_config_text(conf_text, cfg);
    check_supported(cfg);
    check_device(device);
    hdg::FilterConfig fc;
    hdg::parse_filter_text(filter_text, fc);
    *out = new hdg_index{new hdg::Index(cfg, fc, device)};
  });
}

hdg_status hdg_validate_config_text(const char* conf_text, const char* filter_text, char* out, uint64_t cap) {
  return guarded([&] {
    if (!conf_text || !filter_text || !out || cap == 0) hdg::fail(HDG_EINVAL, "null argument");
    hdg::Config cfg;
    hdg::parse_config_text(conf_text, cfg);
    hdg::FilterConfig fc;
    hdg::parse_filter_text(filter_text, fc);
    std::string js = "{";
    static const char* keys[] = {"DATA_WIDTH", "INDEX_WIDTH", "TOPK", "BREAK_DOWN", "ENABLE_MULTI_GPU",
                                 "MEM_BOUND", "IS_CALC_MEM_PREDICTOR_COEFF", "LOWEST_QUERY_BATCH", "FILTER_DIM",
                                 "PQ_DIM", "N_CLUSTERS", "EXPS0", "CLUSTERS", "SUB_CLUSTERS", "MAX_GRIDS", "N_LIST",
                                 "SUB_LISTS", "N_SPLIT", "LABEL_MASK", "IS_SPLIT", "SEED", "KMEANS_ITERS",
                                 "KMEANS_TRAIN", "PQ_TRAIN", "BALANCE_ITERS", "PARITY_BATCH", "QUERY_CHUNK"};
    for (const char* k : keys) {
      uint64_t v = 0;
      hdg::get_config_value(cfg, k, v);
      js += "\"" + std::string(k) + "\": " + std::to_string(v) + ", ";
    }
    std::string path;
    for (char c : cfg.path) {
      if (c == '"' || c == '\\') path += '\\';
      path += c;
    }
    js += "\"PATH\": \"" + path + "\", \"l\": " + std::to_string(fc.l) + ", \"filter_type\": [";
    for (int i = 0; i < fc.l; ++i) js += (i ? ", " : "") + std::to_string(fc.filter_type[i]);
    js += "], \"shift_val\": [";
    for (size_t i = 0; i < fc.shift_val.size(); ++i) {
      char b[64];
      std::snprintf(b, sizeof(b), "%.9g", (double)fc.shift_val[i]);
      js += (i ? ", " : "") + std::string(b);
    }
    js += "], \"div_value\": [";
    for (int i = 0; i < fc.l; ++i) js += (i ? ", " : "") + std::to_string(fc.div_value[i]);
    js += "], \"interval_map\": [";
    for (int i = 0; i < fc.l; ++i) {
      js += i ? ", [" : "[";
      for (size_t j = 0; j < fc.interval_map[i].size(); ++j)
        js += (j ? ", " : "") + std::to_string(fc.interval_map[i][j]);
      js += "]";
    }
    js += "]}";
    if (js.size() + 1 > cap) hdg::fail(HDG_EINVAL, "output buffer too small");
    std::memcpy(out, js.c_str(), js.size() + 1);
  });
}

hdg_status hdg_build(hdg_index* index, const float* dataset, uint64_t n, uint64_t dim, const float* labels,
                     uint64_t label_cols, void* stream) {
  return guarded([&] {
    if (!index) hdg::fail(HDG_EINVAL, "null index");
    index->impl->build(dataset, n, dim, labels, label_cols, (hipStream_t)stream);
  });
}

hdg_status hdg_query(hdg_index* index, const float* queries, uint64_t nq, uint64_t dim, const float* query_labels,
                     uint64_t qlabel_cols, float* out_dist, uint64_t* out_ids, void* stream) {
  return guarded([&] {
    if (!index) hdg::fail(HDG_EINVAL, "null index");
    index->impl->query(queries, nq, dim, query_labels, qlabel_cols, out_dist, out_ids, (hipStream_t)stream);
  });
}

hdg_status hdg_query_noalloc(hdg_index* index, const float* queries, uint64_t nq, uint64_t dim,
                             const float* query_labels, uint64_t qlabel_cols, float* out_dist, uint64_t* out_ids,
                             void* stream) {
  return guarded([&] {
    if (!index) hdg::fail(HDG_EINVAL, "null index");
    index->impl->query_noalloc(queries, nq, dim, query_labels, qlabel_cols, out_dist, out_ids, (hipStream_t)stream);
  });
}

hdg_status hdg_destroy(hdg_index* index) {
  return guarded([&] {
    if (!index) return;
    delete index->impl;
    delete index;
  });
}

hdg_status hdg_set_param(hdg_index* index, const char* key, const char* value) {
  return guarded([&] {
    if (!index || !key || !value) hdg::fail(HDG_EINVAL, "null argument");
    hdg::Config trial = index->impl->config();
    if (!hdg::set_config_value(trial, key, value))
      hdg::fail(HDG_ECONFIG, std::string("unknown HDGrid.conf key: ") + key);
    check_supported(trial);
    index->impl->config() = trial;
  });
}

hdg_status hdg_get_param(const hdg_index* index, const char* key, uint64_t* value) {
  return guarded([&] {
    if (!index || !key || !value) hdg::fail(HDG_EINVAL, "null argument");
    if (!hdg::get_config_value(index->impl->config(), key, *value))
      hdg::fail(HDG_ECONFIG, std::string("unknown integer HDGrid.conf key: ") + key);
  });
}

hdg_status hdg_save(const hdg_index* index, const char* dir) {
  return guarded([&] {
    if (!index || !dir) hdg::fail(HDG_EINVAL, "null argument");
    index->impl->save(dir);
  });
}

hdg_status hdg_load(hdg_index* index, const char* dir, void* stream) {
  return guarded([&] {
    if (!index || !dir) hdg::fail(HDG_EINVAL, "null argument");
    index->impl->load(dir, (hipStream_t)stream);
  });
}

hdg_status hdg_validate_checkpoint(const char* dir, int label_dims) {
  return guarded([&] {
    if (!dir) hdg::fail(HDG_EINVAL, "null argument");
    (void)hdg::read_checkpoint(dir, label_dims);
  });
}

hdg_status hdg_last_query_stats(const hdg_index* index, uint64_t nq, uint64_t* cand, uint64_t* real,
                                uint64_t* refined, uint64_t* scored, uint64_t* adc, uint64_t* kprime) {
  return guarded([&] {
    if (!index) hdg::fail(HDG_EINVAL, "null index");
    uint64_t* const out[6] = {cand, real, refined, scored, adc, kprime};
    index->impl->last_stats(nq, out);
  });
}

hdg_status hdg_last_chosen_subs(const hdg_index* index, uint64_t nq, uint32_t* out) {
  return guarded([&] {
    if (!index || !out) hdg::fail(HDG_EINVAL, "null argument");
    index->impl->last_chosen_subs(nq, out);
  });
}

hdg_status hdg_last_stage_ms(const hdg_index* index, char* names, uint64_t cap, float* ms, uint64_t* n) {
  return guarded([&] {
    if (!index || !n) hdg::fail(HDG_EINVAL, "null argument");
    std::string joined;
    const auto st = index->impl->stage_times();
    for (size_t i = 0; i < st.size(); ++i) {
      if (i) joined += ",";
      joined += st[i].first;
      if (ms) ms[i] = st[i].second;
    }
    *n = st.size();
    if (names && cap) {
      std::strncpy(names, joined.c_str(), cap - 1);
      names[cap - 1] = 0;
    }
  });
}

hdg_status hdg_split_task(const double coeff[4], uint64_t n_data, uint64_t n_queries, uint64_t topk,
                          uint64_t mem_bound, uint64_t available, uint64_t data_width, uint64_t* query_batch,
                          uint64_t* data_batch) {
  return guarded([&] {
    if (!coeff || !query_batch || !data_batch) hdg::fail(HDG_EINVAL, "null argument");
    split_task(coeff, n_data, n_queries, topk, mem_bound, available, data_width, *query_batch, *data_batch);
  });
}

int hdg_find_max_factor(uint64_t a, uint64_t b) { return hdg::find_max_factor(a, b); }

hdg_status hdg_budget_fit(hdg_index* index, const float* queries, uint64_t n, uint64_t dim, const float* query_labels,
                          uint64_t qlabel_cols, const uint64_t* gt_ids, uint64_t k, int levels, double t0, double s0,
                          double ratio, float target, int epochs, float lr, uint64_t seed, float* report,
                          void* stream) {
  return guarded([&] {
    if (!index || !queries || !query_labels || !gt_ids) hdg::fail(HDG_EINVAL, "null argument");
    if (n < 20) hdg::fail(HDG_EINVAL, "budget fit needs at least 20 training queries");
    hdg::BudgetFitParams p;
    p.levels = levels;
    p.t0 = t0;
    p.s0 = s0;
    p.ratio = ratio;
    p.target = target;
    p.epochs = epochs;
    p.lr = lr;
    p.seed = seed;
    index->impl->budget_fit(queries, n, dim, query_labels, qlabel_cols, gt_ids, k, p, (hipStream_t)stream);
    if (report) {
      const hdg::BudgetReport& r = index->impl->budget_report();
      const float vals[8] = {r.final_loss, r.margin, r.holdout_recall, r.holdout_level, r.mean_label,
                             r.fixed_recall, (float)r.n_train, (float)r.n_holdout};
      std::memcpy(report, vals, sizeof(vals));
    }
  });
}

hdg_status hdg_budget_save(const hdg_index* index, const char* path) {
  return guarded([&] {
    if (!index || !path) hdg::fail(HDG_EINVAL, "null argument");
    index->impl->budget_save(path);
  });
}

hdg_status hdg_budget_load(hdg_index* index, const char* path, void* stream) {
  return guarded([&] {
    if (!index || !path) hdg::fail(HDG_EINVAL, "null argument");
    index->impl->budget_load(path, (hipStream_t)stream);
  });
}

hdg_status hdg_budget_debug_grads(const float* params, const float* X, const float* Y, uint32_t n, float* grads,
                                  float* loss) {
  return guarded([&] {
    if (!params || !X || !Y || !grads || !loss || n == 0) hdg::fail(HDG_EINVAL, "null argument");
    hdg::budget_debug_grads(params, X, Y, n, grads, loss);
  });
}

int hdg_budget_param_count(void) { return hdg::budget_param_count(); }

hdg_status hdg_budget_debug_features(hdg_index* index, const float* queries, uint64_t n, uint64_t dim,
                                     const float* query_labels, uint64_t qlabel_cols, float* out, void* stream) {
  return guarded([&] {
    if (!index || !queries || !query_labels || !out) hdg::fail(HDG_EINVAL, "null argument");
    index->impl->budget_debug_features(queries, n, dim, query_labels, qlabel_cols, out, (hipStream_t)stream);
  });
}

hdg_status hdg_debug_coarse_scores(hdg_index* index, const float* queries, uint64_t nq, float* out, void* stream) {
  return guarded([&] {
    if (!index || !queries || !out) hdg::fail(HDG_EINVAL, "null argument");
    index->impl->debug_coarse_scores(queries, nq, out, (hipStream_t)stream);
  });
}

hdg_status hdg_debug_coarse_features(const float* queries, uint64_t nq, uint64_t dim, const uint32_t* list,
                                     uint32_t cap, const uint32_t* n_list, const float* scores, uint64_t ld,
                                     float* feat, int device, void* stream) {
  return guarded([&] {
    if (nq && (!queries || !list || !n_list || !scores || !feat)) hdg::fail(HDG_EINVAL, "null argument");
    if (dim == 0) hdg::fail(HDG_EINVAL, "zero dimension");
    HDG_HIP(hipSetDevice(device));
    hdg::launch_coarse_features(queries, 0, nq, dim, list, cap, n_list, scores, ld, feat, (hipStream_t)stream,
                                nullptr);
    HDG_HIP(hipGetLastError());
  });
}

hdg_status hdg_debug_scored_select(const float* queries, uint64_t nq, uint64_t dim, const uint32_t* list,
                                   uint32_t cap, const uint32_t* n_list, const float* scores, uint64_t ld,
                                   uint32_t k, const uint32_t* budget, uint32_t out_ld, uint32_t* chosen,
                                   uint32_t* n_chosen, float* feat, int wave, int device, void* stream) {
  return guarded([&] {
    if (nq && (!queries || !list || !n_list || !scores || !chosen || !n_chosen)) hdg::fail(HDG_EINVAL, "null argument");
    if (dim == 0 || k == 0 || out_ld < k || cap == 0 || ld < cap) hdg::fail(HDG_EINVAL, "bad sizes");
    if (nq == 0) return;
    HDG_HIP(hipSetDevice(device));
    uint64_t* scratch = nullptr;
    HDG_HIP(hipMalloc((void**)&scratch, nq * cap * sizeof(uint64_t)));
    const hipStream_t s = (hipStream_t)stream;
    hdg::launch_scored_select(queries, 0, nq, dim, list, cap, n_list, scores, ld, k, budget, out_ld, chosen, n_chosen,
                              feat, scratch, wave != 0, s, nullptr, nullptr, 0, nullptr, nullptr, false, false);
    const hipError_t e1 = hipGetLastError();
    const hipError_t e2 = hipStreamSynchronize(s);
    HDG_HIP(hipFree(scratch));
    HDG_HIP(e1);
    HDG_HIP(e2);
  });
}

hdg_status hdg_debug_kmeans_balanced(const float* rows, uint64_t n, uint32_t w, uint32_t k, uint32_t n_iters,
                                     float* centers, uint32_t* labels, int device, void* stream) {
  return guarded([&] {
    if (!rows || !centers || !labels) hdg::fail(HDG_EINVAL, "null argument");
    if (w == 0) hdg::fail(HDG_EINVAL, "zero row width");
    HDG_HIP(hipSetDevice(device));
    std::vector<float> c;
    std::vector<uint32_t> lab;
    hdg::kmeans_balanced(rows, n, w, k, n_iters, c, lab, (hipStream_t)stream);
    std::copy(c.begin(), c.end(), centers);
    std::copy(lab.begin(), lab.end(), labels);
  });
}

hdg_status hdg_merge_topk(const float* dist, const uint64_t* ids, uint64_t n_batches, uint64_t data_batch_size,
                          uint64_t n_data, uint64_t start_offset, uint64_t nq, uint64_t topk, float* out_dist,
                          uint64_t* out_ids, void* stream) {
  return guarded([&] {
    hdg::merge_topk(dist, ids, n_batches, data_batch_size, n_data, start_offset, nq, topk, out_dist, out_ids,
                    (hipStream_t)stream);
  });
}

int hdg_fit_mem_coeff(const double samples[4][3], double coeff[4]) {
  double a[4][5];
  for (int i = 0; i < 4; ++i) {
    a[i][0] = samples[i][0];
    a[i][1] = samples[i][1];
    a[i][2] = samples[i][0] * samples[i][1];
    a[i][3] = 1.0;
    a[i][4] = samples[i][2];
  }
  const int r = gauss(a, 4);
  for (int i = 0; i < 4; ++i) coeff[i] = a[i][4];
  return r;
}

hdg_status hdg_autotune(hdg_index* index, const float* queries, uint64_t nq, uint64_t dim,
                        const float* query_labels, uint64_t qlabel_cols, float* out_dist, uint64_t* out_ids,
                        uint32_t reps, char* report, uint64_t cap, void* stream) {
  return guarded([&] {
    if (!index || !queries || !query_labels || !out_dist || !out_ids) hdg::fail(HDG_EINVAL, "null argument");
    const std::string r = index->impl->autotune(queries, nq, dim, query_labels, qlabel_cols, out_dist, out_ids,
                                                reps, (hipStream_t)stream);
    if (report && cap) {
      std::strncpy(report, r.c_str(), cap - 1);
      report[cap - 1] = 0;
    }
  });
}

}  // extern "C"
