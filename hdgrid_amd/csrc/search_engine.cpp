// search-engine CLI: `search-engine <algorithm> <HDGrid.conf> [filter.conf]`
// (auto-batch-search-main.cu:7-28 + auto_batch_search, search-engine-auto-batch.cu:121-358).
// Reads the dataset directory named by PATH, builds the index on device 0, answers every test query,
// writes res/distances_0 (f32) and res/neighbors_0 (u64) row-major [nq][topk], and reports the recall
// against <PATH>neighbors when that file exists (verification.cuh:97-153, without the
// app_reviews_points side file).
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <unordered_set>
#include <vector>

#include "hdgrid.h"

namespace {

int die(const std::string& msg) {
  std::fprintf(stderr, "search-engine: %s\n", msg.c_str());
  return 1;
}

bool read_lines(const std::string& path, std::vector<std::string>& out) {
  std::ifstream f(path);
  if (!f) return false;
  std::string s;
  while (std::getline(f, s))
    if (!s.empty()) out.push_back(s);
  return true;
}

size_t dtype_size(const std::string& t) {
  static const std::map<std::string, size_t> m = {
      {"int32", 4}, {"float32", 4}, {"float", 4}, {"int64", 8}, {"float64", 8}, {"int8", 1},
      {"uint8", 1}, {"uint16", 2}, {"uint32", 4}, {"uint64", 8}, {"int", 4}};
  auto it = m.find(t);
  return it == m.end() ? 0 : it->second;
}

struct Array {
  std::vector<char> bytes;
  int32_t rows = 0, cols = 0;
  std::string type;
};

bool load_key(const std::string& dir, const std::string& key, const std::string& type, int col_mult, Array& a) {
  std::ifstream fs(dir + key + "_size", std::ios::binary);
  if (!fs) return false;
  int32_t sz[2] = {0, 0};
  fs.read(reinterpret_cast<char*>(sz), 8);
  a.rows = sz[0];
  a.cols = sz[1] * col_mult;
  a.type = type;
  size_t es = dtype_size(type);
  if (!es) return false;
  a.bytes.assign((size_t)a.rows * a.cols * es, 0);
  std::ifstream fd(dir + key, std::ios::binary);
  if (!fd) return false;
  fd.read(a.bytes.data(), a.bytes.size());  // a short file leaves zeros, as io_utils.cu:73-79 does
  return true;
}

// a simple "key = value" lookup on the HDGrid.conf text for the engine's own fields
std::string conf_value(const std::string& text, const std::string& key) {
  std::istringstream in(text);
  std::string line;
  while (std::getline(in, line)) {
    size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    std::string k = line.substr(0, eq), v = line.substr(eq + 1);
    auto trim = [](std::string s) {
      size_t b = s.find_first_not_of(" \t\r"), e = s.find_last_not_of(" \t\r");
      return b == std::string::npos ? std::string() : s.substr(b, e - b + 1);
    };
    k = trim(k);
    for (auto& c : k) c = (char)std::toupper((unsigned char)c);
    if (k == key) {
      v = trim(v);
      if (v.size() >= 2 && v.front() == '"' && v.back() == '"') v = v.substr(1, v.size() - 2);
      return v;
    }
  }
  return "";
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) return die("usage: search-engine <algorithm> <HDGrid.conf> [filter.conf]");
  const std::string algorithm = argv[1], conf_path = argv[2];
  const char* filter_path = argc > 3 ? argv[3] : nullptr;
  std::ifstream cf(conf_path);
  if (!cf) return die("cannot open " + conf_path);
  std::stringstream cs;
  cs << cf.rdbuf();
  const std::string conf = cs.str();

  hdg_index* index = nullptr;
  if (hdg_create(algorithm.c_str(), conf_path.c_str(), filter_path, 0, &index) != HDG_OK)
    return die(hdg_last_error());
  uint64_t topk = 0, filter_dim = 1;
  hdg_get_param(index, "TOPK", &topk);
  hdg_get_param(index, "FILTER_DIM", &filter_dim);
  const std::string dir = conf_value(conf, "PATH");

  std::vector<std::string> keys, types;
  if (!read_lines(dir + "keys", keys)) return die("Failed to open file " + dir + "keys");
  if (!read_lines(dir + "dtypes", types) || types.size() < keys.size()) return die("bad dtypes file");
  std::map<std::string, Array> arr;
  for (size_t i = 0; i < keys.size(); ++i) {
    if (!dtype_size(types[i])) return die("data type not suported: " + types[i]);
    Array a;
    int mult = (keys[i].find("test") != std::string::npos && keys[i].find("label") != std::string::npos)
                   ? (int)filter_dim : 1;
    if (load_key(dir, keys[i], types[i], mult, a)) arr[keys[i]] = std::move(a);
  }
  for (const char* k : {"train_vec", "train_label", "test_vec", "test_label"})
    if (!arr.count(k) || arr[k].type.find("float") != 0) return die(std::string("missing float32 array ") + k);
  const Array &tv = arr["train_vec"], &tl = arr["train_label"], &qv = arr["test_vec"], &ql = arr["test_label"];

  auto to_dev = [](const Array& a) {
    void* p = nullptr;
    if (hipMalloc(&p, a.bytes.size()) != hipSuccess) return (void*)nullptr;
    hipMemcpy(p, a.bytes.data(), a.bytes.size(), hipMemcpyHostToDevice);
    return p;
  };
  float* d_tv = (float*)to_dev(tv);
  float* d_tl = (float*)to_dev(tl);
  float* d_qv = (float*)to_dev(qv);
  float* d_ql = (float*)to_dev(ql);
  if (!d_tv || !d_tl || !d_qv || !d_ql) return die("device allocation failed");
  const uint64_t nq = (uint64_t)qv.rows;
  float* d_dist = nullptr;
  uint64_t* d_ids = nullptr;
  hipMalloc(&d_dist, nq * topk * sizeof(float));
  hipMalloc(&d_ids, nq * topk * sizeof(uint64_t));

  auto t0 = std::chrono::steady_clock::now();
  if (hdg_build(index, d_tv, (uint64_t)tv.rows, (uint64_t)tv.cols, d_tl, (uint64_t)tl.cols, nullptr) != HDG_OK)
    return die(hdg_last_error());
  hipDeviceSynchronize();
  auto t1 = std::chrono::steady_clock::now();
  if (hdg_query(index, d_qv, nq, (uint64_t)qv.cols, d_ql, (uint64_t)ql.cols, d_dist, d_ids, nullptr) != HDG_OK)
    return die(hdg_last_error());
  hipDeviceSynchronize();
  auto t2 = std::chrono::steady_clock::now();

  std::vector<float> dist(nq * topk);
  std::vector<uint64_t> ids(nq * topk);
  hipMemcpy(dist.data(), d_dist, dist.size() * sizeof(float), hipMemcpyDeviceToHost);
  hipMemcpy(ids.data(), d_ids, ids.size() * sizeof(uint64_t), hipMemcpyDeviceToHost);
  ::mkdir("res", 0755);
  std::ofstream("res/distances_0", std::ios::binary).write((const char*)dist.data(), dist.size() * sizeof(float));
  std::ofstream("res/neighbors_0", std::ios::binary).write((const char*)ids.data(), ids.size() * sizeof(uint64_t));
  const double build_s = std::chrono::duration<double>(t1 - t0).count();
  const double query_s = std::chrono::duration<double>(t2 - t1).count();
  std::printf("build time: %.3f s, query time: %.3f s, QPS: %.1f\n", build_s, query_s, nq / query_s);

  std::ifstream gt(dir + "neighbors", std::ios::binary);
  if (gt) {
    std::vector<int32_t> g(nq * topk, -1);
    gt.read(reinterpret_cast<char*>(g.data()), g.size() * sizeof(int32_t));
    uint64_t hits = 0, valid = 0;
    for (uint64_t q = 0; q < nq; ++q) {
      std::unordered_set<int64_t> gs, seen;
      for (uint64_t j = 0; j < topk; ++j) {
        int32_t x = g[q * topk + j];
        if (x >= 0 && gs.insert(x).second) ++valid;
      }
      for (uint64_t j = 0; j < topk; ++j) {
        int64_t x = (int64_t)ids[q * topk + j];
        if (!seen.insert(x).second) continue;
        hits += gs.count(x);
      }
    }
    std::printf("final recall: %.6f\n", valid ? (double)hits / valid : 0.0);
  }
  hipFree(d_tv);
  hipFree(d_tl);
  hipFree(d_qv);
  hipFree(d_ql);
  hipFree(d_dist);
  hipFree(d_ids);
  hdg_destroy(index);
  return 0;
}
