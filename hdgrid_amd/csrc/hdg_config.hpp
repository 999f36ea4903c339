// HDGrid.conf and filter.conf parsers with the reference's grammar.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace hdg {

// Keys of the reference's HDGrid.conf (config_utils.h:97-108, product-quantization.cuh:32-36,
// Super-Grid-IVF_PQ.cuh:46-55).  Missing keys keep the defaults below (the README sample values);
// the reference leaves them uninitialised.  Build-side keys that the reference does not have are
// listed after the reference keys; the reference parser logs and skips them.
struct Config {
  // common_config
  uint64_t data_width = 1;
  uint64_t index_width = 0;
  uint64_t topk = 100;
  uint64_t break_down = 1;
  uint64_t enable_multi_gpu = 0;
  uint64_t mem_bound = 5000000000ull;
  uint64_t is_calc_mem_predictor_coeff = 0;
  uint64_t lowest_query_batch = 125;
  uint64_t filter_dim = 1;
  std::string path;
  // pq::index_params
  uint64_t pq_dim = 4;
  uint64_t n_clusters = 256;
  uint64_t exps0 = 30;
  // super_grid::index_params
  uint64_t clusters = 512;
  uint64_t sub_clusters = 10;
  uint64_t max_grids = 30000000;
  uint64_t n_list = 64;
  uint64_t sub_lists = 128;
  uint64_t n_split = 1;
  uint64_t label_mask = 0;
  uint64_t is_split = 0;
  // build-side keys of this implementation
  uint64_t seed = 0x4844ull;           // SEED: k-means / sampling seed
  uint64_t kmeans_iters = 16;          // KMEANS_ITERS: Lloyd iterations (fine + PQ)
  uint64_t kmeans_train = 262144;      // KMEANS_TRAIN: max training rows for the fine k-means
  uint64_t pq_train = 65536;           // PQ_TRAIN: max training rows per PQ subspace
  uint64_t balance_iters = 200;        // BALANCE_ITERS: balanced k-means iterations of the grouping
  uint64_t parity_batch = 0;           // PARITY_BATCH: beta group size (0 = findMaxFactor(125, nq))
  uint64_t query_chunk = 16384;        // QUERY_CHUNK: queries processed per device pass
  uint64_t precision = 0;              // PRECISION: 0 reference-exact FP32, 1 BF16 coarse + FP16 LUT
  uint64_t probe_clusters = 0;         // PROBE_CLUSTERS: clusters probed per query in PRECISION 1 (0 = N_LIST)
  uint64_t budget_model = 0;           // BUDGET_MODEL: 1 = per-query budgets from the loaded MLP
  uint64_t sparse_count = 1;           // SPARSE_COUNT: 1 = bucket-grid walk for selective boxes, 0 = dense count
  uint64_t sparse_max_tests = 131072;  // SPARSE_MAX_TESTS: entries a query may test on the walk before the dense count
  uint64_t adc_per = 32;               // ADC_PER: candidates per thread of an ADC chunk (8, 16, 32, 64)
  uint64_t adc_wgs = 2048;             // ADC_WGS: persistent ADC workgroups
  uint64_t graph = 0;                  // GRAPH: 1 = replay each query chunk as a captured HIP graph
  uint64_t select_wave = 1;            // SELECT_WAVE: 1 = one wave per query for scored selection, in-box keep
                                       // and the final top-k on lists of <= 512 items (same results)
  uint64_t score_block = 65536;        // SCORE_BLOCK: queries per BF16 scoring block (PRECISION 1; min 128)
  uint64_t adc_split = 65536;          // ADC_SPLIT: ADC candidates from which a query's top-k' selection runs
                                       // over many workgroups (0 = never; same results)
  uint64_t inbox_cap = 4096;           // INBOX_CAP: in-box points the bucket-grid walk lists per query (512..2^20);
                                       // under PRECISION 1 a walked query whose list fits refines from it
  uint64_t budget_labels = 0;          // BUDGET_LABELS: path whose traces label the budget model's training
                                       // queries: 0 = plain (PRECISION 0), 1 = mixed (PRECISION 1)
  uint64_t label_words = 1;            // LABEL_WORDS: 1 = label post-filters read the packed label words when the
                                       // index has them (integer labels whose ranges fit; same results)
};

// Parse with config_utils.h:33-88 semantics: `KEY = value` per line, key case-insensitive and
// trimmed; integer values stored as u64; other unquoted numbers written as f32 bits into the low
// 4 bytes of the field; quoted values only for PATH.  Unknown keys are logged and skipped.
// Throws hdg::Error(kConfigError) for an unreadable file or a value the reference would crash on.
void parse_config_file(const std::string& path, Config& cfg);
void parse_config_text(const std::string& text, Config& cfg);
// Set one key (same value grammar); returns false for an unknown key.
bool set_config_value(Config& cfg, const std::string& key, const std::string& value);
// Read an integer key (case-insensitive); returns false for PATH or an unknown key.
bool get_config_value(const Config& cfg, const std::string& key, uint64_t& value);

// filter.conf (config_utils.h:112-172).
struct FilterConfig {
  int l = 0;
  std::vector<int> filter_type;               // [l]
  std::vector<float> shift_val;               // [2l]
  std::vector<int> div_value;                 // [l]
  std::vector<std::vector<int>> interval_map; // [l]
  // all interval maps concatenated in filter order, as f32 (preprocessing.cu:145-150)
  std::vector<float> flat_map() const;
};
void parse_filter_file(const std::string& path, FilterConfig& fc);
void parse_filter_text(const std::string& text, FilterConfig& fc);

}  // namespace hdg
