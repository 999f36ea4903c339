// Test-infrastructure driver around the reference's own host-only code, compiled from the reference
// tree (src/utils/config_utils.h + string_utils.h + the vendored easylogging++, src/utils/math_utils.h).
// It prints what the reference parsers and helpers produce so that tests/golden/ can hold those outputs
// as fixtures.  Nothing here is shipped or linked into the product.
//
//   ref_host_dump conf <HDGrid.conf>        -> JSON of every HDGrid.conf field (u64 bit patterns)
//   ref_host_dump filter <filter.conf>      -> JSON of the parsed filter_config
//   ref_host_dump factor <a> <b>            -> findMaxFactor(a, b)
//   ref_host_dump divide <n> <m> <i>        -> uniformly_divide(n, m, i): size offset
//   ref_host_dump gauss <d q used> x 4      -> gauss() on calc_mem_coeff's system
//                                              (search-engine-auto-batch.cu:404-427): code c0 c1 c2 c3
#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
using std::max;

#include "utils/config_utils.h"
#include "utils/math_utils.h"

INITIALIZE_EASYLOGGINGPP

namespace {

// Field block laid out like pq::index_params + super_grid::index_params, registered through the
// reference's extender mechanism (the keys of product-quantization.cuh:32-36, Super-Grid-IVF_PQ.cuh:46-55).
struct dump_config : common_config {
  uint64_t pq_dim, n_clusters, exps0, pq_len;
  uint64_t clusters, sub_clusters, max_grids, n_list, sub_lists, n_split, label_mask, is_split;

  explicit dump_config(const std::string& p)
      : common_config(p, [](std::map<std::string, int>& m) {
          m.insert({{"PQ_DIM", OFFSET_OF(dump_config, pq_dim)},
                    {"N_CLUSTERS", OFFSET_OF(dump_config, n_clusters)},
                    {"EXPS0", OFFSET_OF(dump_config, exps0)},
                    {"CLUSTERS", OFFSET_OF(dump_config, clusters)},
                    {"SUB_CLUSTERS", OFFSET_OF(dump_config, sub_clusters)},
                    {"MAX_GRIDS", OFFSET_OF(dump_config, max_grids)},
                    {"N_LIST", OFFSET_OF(dump_config, n_list)},
                    {"SUB_LISTS", OFFSET_OF(dump_config, sub_lists)},
                    {"N_SPLIT", OFFSET_OF(dump_config, n_split)},
                    {"LABEL_MASK", OFFSET_OF(dump_config, label_mask)},
                    {"IS_SPLIT", OFFSET_OF(dump_config, is_split)}});
        }) {
    data_width = index_width = topk = break_down = enable_multi_gpu = mem_bound = 0;
    is_calc_mem_predictor_coeff = lowest_query_batch = filter_dim = 0;
    pq_dim = n_clusters = exps0 = pq_len = 0;
    clusters = sub_clusters = max_grids = n_list = sub_lists = n_split = label_mask = is_split = 0;
  }
};

void dump_conf(const char* path) {
  dump_config c(path);
  c.read_config();
  std::cout << "{\"DATA_WIDTH\": " << c.data_width << ", \"INDEX_WIDTH\": " << c.index_width
            << ", \"TOPK\": " << c.topk << ", \"BREAK_DOWN\": " << c.break_down
            << ", \"ENABLE_MULTI_GPU\": " << c.enable_multi_gpu << ", \"MEM_BOUND\": " << c.mem_bound
            << ", \"IS_CALC_MEM_PREDICTOR_COEFF\": " << c.is_calc_mem_predictor_coeff
            << ", \"LOWEST_QUERY_BATCH\": " << c.lowest_query_batch << ", \"FILTER_DIM\": " << c.filter_dim
            << ", \"PQ_DIM\": " << c.pq_dim << ", \"N_CLUSTERS\": " << c.n_clusters << ", \"EXPS0\": " << c.exps0
            << ", \"CLUSTERS\": " << c.clusters << ", \"SUB_CLUSTERS\": " << c.sub_clusters
            << ", \"MAX_GRIDS\": " << c.max_grids << ", \"N_LIST\": " << c.n_list << ", \"SUB_LISTS\": " << c.sub_lists
            << ", \"N_SPLIT\": " << c.n_split << ", \"LABEL_MASK\": " << c.label_mask << ", \"IS_SPLIT\": " << c.is_split
            << ", \"PATH\": \"" << c.path << "\"}" << std::endl;
}

void dump_filter(const char* path) {
  filter_config f(path);
  std::cout << "{\"l\": " << f.l << ", \"filter_type\": [";
  for (int i = 0; i < f.l; ++i) std::cout << (i ? ", " : "") << f.filter_type[i];
  std::cout << "], \"shift_val\": [";
  for (size_t i = 0; i < f.shift_val.size(); ++i) {
    char b[64];
    std::snprintf(b, sizeof(b), "%.9g", (double)f.shift_val[i]);
    std::cout << (i ? ", " : "") << b;
  }
  std::cout << "], \"div_value\": [";
  for (int i = 0; i < f.l; ++i) std::cout << (i ? ", " : "") << f.div_value[i];
  std::cout << "], \"interval_map\": [";
  for (int i = 0; i < f.l; ++i) {
    std::cout << (i ? ", [" : "[");
    for (size_t j = 0; j < f.interval_map[i].size(); ++j) std::cout << (j ? ", " : "") << f.interval_map[i][j];
    std::cout << "]";
  }
  std::cout << "]}" << std::endl;
}

}  // namespace

int main(int argc, char** argv) {
  el::Configurations conf;
  conf.setToDefault();
  conf.set(el::Level::Global, el::ConfigurationType::Enabled, "false");
  el::Loggers::reconfigureAllLoggers(conf);
  if (argc >= 3 && !std::strcmp(argv[1], "conf")) {
    dump_conf(argv[2]);
  } else if (argc >= 3 && !std::strcmp(argv[1], "filter")) {
    dump_filter(argv[2]);
  } else if (argc >= 4 && !std::strcmp(argv[1], "factor")) {
    std::cout << findMaxFactor(std::stoull(argv[2]), std::stoull(argv[3])) << std::endl;
  } else if (argc >= 5 && !std::strcmp(argv[1], "divide")) {
    uint64_t size = 0, offset = 0;
    uniformly_divide(std::stoull(argv[2]), std::stoull(argv[3]), std::stoull(argv[4]), size, offset);
    std::cout << size << " " << offset << std::endl;
  } else if (argc >= 14 && !std::strcmp(argv[1], "gauss")) {
    double a[4][5];
    for (int i = 0; i < 4; ++i) {
      const double d = std::stod(argv[2 + 3 * i]), q = std::stod(argv[3 + 3 * i]), used = std::stod(argv[4 + 3 * i]);
      a[i][0] = d;
      a[i][1] = q;
      a[i][2] = d * q;
      a[i][3] = 1;
      a[i][4] = used;
    }
    const int r = gauss(a, 4);
    char b[128];
    std::snprintf(b, sizeof(b), "%d %.17g %.17g %.17g %.17g", r, a[0][4], a[1][4], a[2][4], a[3][4]);
    std::cout << b << std::endl;
  } else {
    std::cerr << "usage: ref_host_dump conf|filter <path> | factor <a> <b> | divide <n> <m> <i> | gauss <12 numbers>" << std::endl;
    return 2;
  }
  return 0;
}
