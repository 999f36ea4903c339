// HDGrid.conf / filter.conf parsing.  Grammar restated from config_utils.h:33-88 (HDGrid.conf),
// config_utils.h:121-172 (filter.conf) and string_utils.h:46-97 (trim, isInteger, parse_array).
#include "hdg_config.hpp"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "hdg_common.hpp"

namespace hdg {

int log_level() {
  static int lvl = [] {
    const char* e = std::getenv("HDG_LOG");
    return e ? std::atoi(e) : 0;
  }();
  return lvl;
}

void log_line(int level, const char* fmt, ...) {
  if (level > log_level()) return;
  static const char* names[] = {"ERROR", "INFO", "TRACE"};
  std::fprintf(stderr, "[hdgrid %s] ", names[level < 0 ? 0 : (level > 2 ? 2 : level)]);
  va_list ap;
  va_start(ap, fmt);
  std::vfprintf(stderr, fmt, ap);
  va_end(ap);
  std::fputc('\n', stderr);
}

namespace {

std::string trim_ws(std::string s) {
  size_t b = 0;
  while (b < s.size() && std::isspace((unsigned char)s[b])) ++b;
  size_t e = s.size();
  while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
  return s.substr(b, e - b);
}

std::string upper(std::string s) {
  for (char& c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}

// string_utils.h:71-85: optional sign then digits only; a lone sign counts as an integer.
bool is_integer(const std::string& s) {
  if (s.empty()) return false;
  size_t i = 0;
  if (s[0] == '+' || s[0] == '-') i = 1;
  for (; i < s.size(); ++i)
    if (!std::isdigit((unsigned char)s[i])) return false;
  return i > 0;
}

std::string strip_spaces(const std::string& s) {
  std::string r;
  for (char c : s)
    if (!std::isspace((unsigned char)c)) r += c;
  return r;
}

// string_utils.h:100-124 (parse_array): optional [ ], comma-separated, each token via stod.
template <class T>
std::vector<T> parse_list(const std::string& str) {
  std::string s = str;
  if (!s.empty() && s.front() == '[') s.erase(0, 1);
  if (!s.empty() && s.back() == ']') s.pop_back();
  std::vector<T> out;
  std::stringstream ss(s);
  std::string tok;
  while (std::getline(ss, tok, ',')) {
    size_t b = tok.find_first_not_of(" \t");
    if (b == std::string::npos) continue;
    size_t e = tok.find_last_not_of(" \t");
    tok = tok.substr(b, e - b + 1);
    char* end = nullptr;
    errno = 0;
    double v = std::strtod(tok.c_str(), &end);
    if (end == tok.c_str()) fail(kConfigError, "filter.conf: not a number: '" + tok + "'");
    out.push_back(static_cast<T>(v));
  }
  return out;
}

struct FieldRef {
  uint64_t* u = nullptr;
  std::string* s = nullptr;
};

FieldRef lookup(Config& c, const std::string& key) {
  static const char* const names[] = {
      "DATA_WIDTH", "INDEX_WIDTH", "TOPK", "BREAK_DOWN", "ENABLE_MULTI_GPU", "MEM_BOUND",
      "IS_CALC_MEM_PREDICTOR_COEFF", "LOWEST_QUERY_BATCH", "FILTER_DIM", "PQ_DIM", "N_CLUSTERS",
      "EXPS0", "CLUSTERS", "SUB_CLUSTERS", "MAX_GRIDS", "N_LIST", "SUB_LISTS", "N_SPLIT",
      "LABEL_MASK", "IS_SPLIT", "SEED", "KMEANS_ITERS", "KMEANS_TRAIN", "PQ_TRAIN", "BALANCE_ITERS",
      "PARITY_BATCH", "QUERY_CHUNK", "PRECISION", "PROBE_CLUSTERS", "BUDGET_MODEL", "SPARSE_COUNT",
      "SPARSE_MAX_TESTS", "ADC_PER", "ADC_WGS", "GRAPH", "SELECT_WAVE", "SCORE_BLOCK", "ADC_SPLIT",
      "INBOX_CAP", "BUDGET_LABELS", "LABEL_WORDS"};
  uint64_t* const fields[] = {
      &c.data_width, &c.index_width, &c.topk, &c.break_down, &c.enable_multi_gpu, &c.mem_bound,
      &c.is_calc_mem_predictor_coeff, &c.lowest_query_batch, &c.filter_dim, &c.pq_dim,
      &c.n_clusters, &c.exps0, &c.clusters, &c.sub_clusters, &c.max_grids, &c.n_list, &c.sub_lists,
      &c.n_split, &c.label_mask, &c.is_split, &c.seed, &c.kmeans_iters, &c.kmeans_train,
      &c.pq_train, &c.balance_iters, &c.parity_batch, &c.query_chunk, &c.precision, &c.probe_clusters,
      &c.budget_model, &c.sparse_count, &c.sparse_max_tests, &c.adc_per, &c.adc_wgs, &c.graph, &c.select_wave, &c.score_block, &c.adc_split,
      &c.inbox_cap, &c.budget_labels, &c.label_words};
  FieldRef r;
  if (key == "PATH") {
    r.s = &c.path;
    return r;
  }
  for (size_t i = 0; i < sizeof(names) / sizeof(names[0]); ++i)
    if (key == names[i]) r.u = fields[i];
  return r;
}

// Apply one `key = value` line.  Returns false when the key is unknown.
bool apply_line(Config& cfg, const std::string& line) {
  size_t eq = line.find('=');
  std::string key = trim_ws(upper(line.substr(0, eq)));
  std::string val = trim_ws(line.substr(eq == std::string::npos ? 0 : eq + 1));
  size_t qs = val.find('"');
  size_t qe = std::string::npos;
  if (qs != std::string::npos) {
    qe = val.find('"', qs + 1);
    while (qe != std::string::npos && val.at(qe - 1) == '\\') {
      val.erase(qe - 1, 1);
      qe = val.find('"', qe + 2);
    }
  }
  if (qs != std::string::npos && qe != std::string::npos && qs != qe && qs + 1 != qe)
    val = val.substr(qs + 1, qe - 1);
  FieldRef f = lookup(cfg, key);
  if (!f.u && !f.s) {
    if (!key.empty()) log_line(0, "ERROR: Key '%s' not found in str_to_offset_map!", key.c_str());
    return false;
  }
  if (qs == std::string::npos) {
    if (!f.u) fail(kConfigError, "HDGrid.conf: PATH must be quoted");
    if (is_integer(val)) {
      *f.u = (uint64_t)std::atoll(val.c_str());
    } else {
      char* end = nullptr;
      errno = 0;
      float v = std::strtof(val.c_str(), &end);
      if (end == val.c_str() || errno == ERANGE)
        fail(kConfigError, "HDGrid.conf: value of " + key + " is not a number: '" + val + "'");
      // the reference stores the f32 bits into the first 4 bytes of the u64 field
      uint32_t bits;
      std::memcpy(&bits, &v, 4);
      *f.u = (*f.u & 0xFFFFFFFF00000000ull) | bits;
      log_line(0, "HDGrid.conf: non-integer value for %s stored as f32 bits (reference behaviour)",
               key.c_str());
    }
  } else {
    if (!f.s) fail(kConfigError, "HDGrid.conf: quoted value for numeric key " + key);
    *f.s = val;
  }
  return true;
}

}  // namespace

void parse_config_text(const std::string& text, Config& cfg) {
  std::istringstream in(text);
  std::string line;
  while (in.good()) {
    std::getline(in, line);
    apply_line(cfg, line);
  }
}

void parse_config_file(const std::string& path, Config& cfg) {
  std::ifstream f(path);
  if (!f.is_open()) fail(kConfigError, "cannot open config file: " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  parse_config_text(ss.str(), cfg);
}

bool set_config_value(Config& cfg, const std::string& key, const std::string& value) {
  return apply_line(cfg, key + " = " + value);
}

bool get_config_value(const Config& cfg, const std::string& key, uint64_t& value) {
  FieldRef f = lookup(const_cast<Config&>(cfg), trim_ws(upper(key)));
  if (!f.u) return false;
  value = *f.u;
  return true;
}

std::vector<float> FilterConfig::flat_map() const {
  std::vector<float> out;
  for (const auto& m : interval_map)
    for (int v : m) out.push_back((float)v);
  return out;
}

void parse_filter_text(const std::string& text, FilterConfig& fc) {
  std::istringstream in(text);
  std::string line;
  auto check_idx = [&](int idx) {
    if (idx < 0 || idx >= fc.l) fail(kConfigError, "filter.conf: filter index outside [0, l)");
  };
  while (std::getline(in, line)) {
    line = strip_spaces(line);
    if (line.empty() || line[0] == '#') continue;
    if (line.rfind("l=", 0) == 0) {
      fc.l = std::atoi(line.substr(2).c_str());
      if (fc.l < 1 || fc.l > 8) fail(kConfigError, "filter.conf: l must be in [1, 8]");
      fc.filter_type.resize(fc.l);
      fc.shift_val.resize(2 * fc.l);
      fc.interval_map.resize(fc.l);
      fc.div_value.resize(fc.l);
    } else if (line.rfind("filter", 0) == 0) {
      int idx = std::atoi(line.substr(6).c_str());
      check_idx(idx);
      while (std::getline(in, line)) {
        line = strip_spaces(line);
        if (line.empty()) break;
        if (line.rfind("type=", 0) == 0) {
          int t = std::atoi(line.substr(5).c_str());
          fc.filter_type[idx] = t;
          if (t == 3) break;
        } else if (line.rfind("shift_val=", 0) == 0) {
          auto v = parse_list<float>(line.substr(10));
          if (v.size() < 2) fail(kConfigError, "filter.conf: shift_val needs two values");
          fc.shift_val[2 * idx] = v[0];
          fc.shift_val[2 * idx + 1] = v[1];
          break;
        } else if (line.rfind("interval_map=", 0) == 0) {
          fc.interval_map[idx] = parse_list<int>(line.substr(13));
          break;
        } else if (line.rfind("div_value=", 0) == 0) {
          auto v = parse_list<int>(line.substr(10));
          if (v.empty()) fail(kConfigError, "filter.conf: div_value needs a value");
          fc.div_value[idx] = v[0];
          break;
        }
      }
    }
  }
  if (fc.l < 1) fail(kConfigError, "filter.conf: missing `l = <dims>` line");
  for (int i = 0; i < fc.l; ++i) {
    int t = fc.filter_type[i];
    if (t < 0 || t > 3) fail(kConfigError, "filter.conf: filter type must be 0..3");
    if (t == 2 && fc.div_value[i] == 0) fail(kConfigError, "filter.conf: div_value must be non-zero");
  }
}

void parse_filter_file(const std::string& path, FilterConfig& fc) {
  std::ifstream f(path);
  if (!f.is_open()) fail(kConfigError, "Unable to open configuration file: " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  parse_filter_text(ss.str(), fc);
}

}  // namespace hdg
