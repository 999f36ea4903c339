// search-engine CLI: `search-engine <algorithm> <HDGrid.conf> [filter.conf]`
// (auto-batch-search-main.cu:7-28 + auto_batch_search, search-engine-auto-batch.cu:121-358).
//
// Reads the dataset directory named by PATH and runs the engine loop on device 0:
//   1. memory predictor: coefficients from ./coeff (4 f64), refitted from four probe runs when
//      IS_CALC_MEM_PREDICTOR_COEFF = 1 (calc_mem_coeff, search-engine-auto-batch.cu:381-428), then
//      split_task -> (data batch, query batch) under min(free/5, MEM_BOUND) (:430-489);
//   2. per data batch: build the index on its rows, answer every test query in blocks of whole query
//      batches (the query batch is the beta group, PARITY_BATCH), each block's results copied to pinned host
//      memory on a side stream while the next block computes (flush_current_res, :20-103);
//   3. more than one data batch: the per-batch results stay in device memory and are merged on the device
//      with integer id offsets (merge_intermediate_result, merge_res.cu:18-55);
//   4. writes res/distances_0 (f32) and res/neighbors_0 (u64) row-major [nq][topk] (replacing earlier
//      files) and reports the recall against <PATH>neighbors when that file exists
//      (verification.cuh:97-153, without the app_reviews_points side file).
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <algorithm>
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


bool read_coeff(const std::string& path, double coeff[4]) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  f.read(reinterpret_cast<char*>(coeff), 4 * sizeof(double));
  return (bool)f;
}

bool write_coeff(const std::string& path, const double coeff[4]) {
  std::ofstream f(path, std::ios::binary);
  if (!f) return false;
  f.write(reinterpret_cast<const char*>(coeff), 4 * sizeof(double));
  return (bool)f;
}

#define SE_HIP(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) return die(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct DevArrays {
  float *tv = nullptr, *tl = nullptr, *qv = nullptr, *ql = nullptr, *dist = nullptr;
  uint64_t* ids = nullptr;
  float* mdist = nullptr;
  uint64_t* mids = nullptr;
  ~DevArrays() {
    for (void* p : {(void*)tv, (void*)tl, (void*)qv, (void*)ql, (void*)dist, (void*)ids, (void*)mdist, (void*)mids})
      if (p) (void)hipFree(p);
  }
};

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3 || argc > 4) return die("usage: search-engine <algorithm> <HDGrid.conf> [filter.conf]");
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
  uint64_t topk = 0, filter_dim = 1, mem_bound = 0, data_width = 1, calc_coeff = 0, break_down = 0;
  hdg_get_param(index, "TOPK", &topk);
  hdg_get_param(index, "FILTER_DIM", &filter_dim);
  hdg_get_param(index, "MEM_BOUND", &mem_bound);
  hdg_get_param(index, "DATA_WIDTH", &data_width);
  hdg_get_param(index, "IS_CALC_MEM_PREDICTOR_COEFF", &calc_coeff);
  hdg_get_param(index, "BREAK_DOWN", &break_down);
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
  const uint64_t n_data = (uint64_t)tv.rows, dim = (uint64_t)tv.cols, nq = (uint64_t)qv.rows;
  const uint64_t lcols = (uint64_t)tl.cols, qcols = (uint64_t)ql.cols;
  if (n_data == 0 || nq == 0) return die("empty train_vec or test_vec");

  DevArrays dv;
  // ---- 1. memory predictor ----
  double coeff[4] = {0, 0, 0, 0};
  if (calc_coeff) {
    // four probe runs (calc_mem_coeff's sizes), device bytes in use measured around each
    const uint64_t probe_d[4] = {10000, 20000, 30000, 40000}, probe_q[4] = {5, 10, 25, 20};
    double samples[4][3];
    for (int i = 0; i < 4; ++i) {
      const uint64_t pd = std::min(probe_d[i], n_data), pq = std::min(probe_q[i], nq);
      SE_HIP(hipDeviceSynchronize());
      size_t free0 = 0, total = 0;
      SE_HIP(hipMemGetInfo(&free0, &total));
      hdg_index* probe = nullptr;
      if (hdg_create(algorithm.c_str(), conf_path.c_str(), filter_path, 0, &probe) != HDG_OK)
        return die(hdg_last_error());
      void *p_tv = nullptr, *p_tl = nullptr, *p_qv = nullptr, *p_ql = nullptr, *p_d = nullptr, *p_i = nullptr;
      SE_HIP(hipMalloc(&p_tv, pd * dim * 4));
      SE_HIP(hipMalloc(&p_tl, pd * lcols * 4));
      SE_HIP(hipMalloc(&p_qv, pq * dim * 4));
      SE_HIP(hipMalloc(&p_ql, pq * qcols * 4));
      SE_HIP(hipMalloc(&p_d, pq * topk * 4));
      SE_HIP(hipMalloc(&p_i, pq * topk * 8));
      SE_HIP(hipMemcpy(p_tv, tv.bytes.data(), pd * dim * 4, hipMemcpyHostToDevice));
      SE_HIP(hipMemcpy(p_tl, tl.bytes.data(), pd * lcols * 4, hipMemcpyHostToDevice));
      SE_HIP(hipMemcpy(p_qv, qv.bytes.data(), pq * dim * 4, hipMemcpyHostToDevice));
      SE_HIP(hipMemcpy(p_ql, ql.bytes.data(), pq * qcols * 4, hipMemcpyHostToDevice));
      if (hdg_build(probe, (float*)p_tv, pd, dim, (float*)p_tl, lcols, nullptr) != HDG_OK ||
          hdg_query(probe, (float*)p_qv, pq, dim, (float*)p_ql, qcols, (float*)p_d, (uint64_t*)p_i, nullptr) !=
              HDG_OK)
        return die(hdg_last_error());
      SE_HIP(hipDeviceSynchronize());
      size_t free1 = 0;
      SE_HIP(hipMemGetInfo(&free1, &total));
      samples[i][0] = (double)pd;
      samples[i][1] = (double)pq;
      samples[i][2] = (double)(free0 > free1 ? free0 - free1 : 0);
      hdg_destroy(probe);
      for (void* p : {p_tv, p_tl, p_qv, p_ql, p_d, p_i}) (void)hipFree(p);
    }
    const int g = hdg_fit_mem_coeff(samples, coeff);
    if (g != 0) std::fprintf(stderr, "search-engine: predictor fit is %s\n", g == 1 ? "rank deficient" : "inconsistent");
    if (!write_coeff("coeff", coeff)) return die("Failed to open file for writing: coeff");
  } else if (!read_coeff("coeff", coeff)) {
    // no ./coeff: a model of this library's own footprint -- per data row the vector, PQ codes, labels
    // and the z-pool / bucket-grid entries; per query its workspace; a fixed 256 MiB
    uint64_t pq_dim = 0;
    hdg_get_param(index, "PQ_DIM", &pq_dim);
    coeff[0] = 4.0 * (double)dim + (double)pq_dim + 8.0 * (double)lcols + 64.0;
    coeff[1] = 1 << 20;
    coeff[2] = 0.0;
    coeff[3] = 256.0 * (1 << 20);
    std::fprintf(stderr, "search-engine: no ./coeff, using the built-in footprint model\n");
  }
  size_t free_b = 0, total_b = 0;
  SE_HIP(hipMemGetInfo(&free_b, &total_b));
  uint64_t qb = 0, db = 0;
  if (hdg_split_task(coeff, n_data, nq, topk, mem_bound, free_b, data_width, &qb, &db) != HDG_OK)
    return die(hdg_last_error());
  const uint64_t n_batches = (n_data + db - 1) / db;
  std::printf("device: 0, data batch size: %llu, query batch size: %llu, data batches: %llu\n",
              (unsigned long long)db, (unsigned long long)qb, (unsigned long long)n_batches);
  if (hdg_set_param(index, "PARITY_BATCH", std::to_string(qb).c_str()) != HDG_OK) return die(hdg_last_error());

  // ---- 2. per data batch: build, then the queries in blocks with overlapped copies ----
  SE_HIP(hipMalloc((void**)&dv.tv, db * dim * 4));
  SE_HIP(hipMalloc((void**)&dv.tl, db * lcols * 4));
  SE_HIP(hipMalloc((void**)&dv.qv, nq * dim * 4));
  SE_HIP(hipMalloc((void**)&dv.ql, nq * qcols * 4));
  SE_HIP(hipMemcpy(dv.qv, qv.bytes.data(), nq * dim * 4, hipMemcpyHostToDevice));
  SE_HIP(hipMemcpy(dv.ql, ql.bytes.data(), nq * qcols * 4, hipMemcpyHostToDevice));
  const uint64_t res_n = nq * topk;
  SE_HIP(hipMalloc((void**)&dv.dist, n_batches * res_n * 4));
  SE_HIP(hipMalloc((void**)&dv.ids, n_batches * res_n * 8));
  float* h_dist = nullptr;
  uint64_t* h_ids = nullptr;
  SE_HIP(hipHostMalloc((void**)&h_dist, res_n * 4, hipHostMallocDefault));
  SE_HIP(hipHostMalloc((void**)&h_ids, res_n * 8, hipHostMallocDefault));
  hipStream_t comp = nullptr, copy = nullptr;
  SE_HIP(hipStreamCreateWithFlags(&comp, hipStreamNonBlocking));
  SE_HIP(hipStreamCreateWithFlags(&copy, hipStreamNonBlocking));
  const uint64_t block = qb * std::max<uint64_t>(1, 16384 / qb);
  std::vector<hipEvent_t> done((nq + block - 1) / block);
  for (auto& e : done) SE_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));

  double build_s = 0.0, query_s = 0.0;
  for (uint64_t b = 0; b < n_batches; ++b) {
    const uint64_t off = b * db, rows = std::min(db, n_data - off);
    SE_HIP(hipMemcpy(dv.tv, tv.bytes.data() + off * dim * 4, rows * dim * 4, hipMemcpyHostToDevice));
    SE_HIP(hipMemcpy(dv.tl, tl.bytes.data() + off * lcols * 4, rows * lcols * 4, hipMemcpyHostToDevice));
    auto t0 = std::chrono::steady_clock::now();
    if (hdg_build(index, dv.tv, rows, dim, dv.tl, lcols, comp) != HDG_OK) return die(hdg_last_error());
    SE_HIP(hipStreamSynchronize(comp));
    auto t1 = std::chrono::steady_clock::now();
    build_s += std::chrono::duration<double>(t1 - t0).count();
    float* bd = dv.dist + b * res_n;
    uint64_t* bi = dv.ids + b * res_n;
    for (uint64_t q0 = 0, bk = 0; q0 < nq; q0 += block, ++bk) {
      const uint64_t qn = std::min(block, nq - q0);
      if (hdg_query(index, dv.qv + q0 * dim, qn, dim, dv.ql + q0 * qcols, qcols, bd + q0 * topk, bi + q0 * topk,
                    comp) != HDG_OK)
        return die(hdg_last_error());
      if (break_down) {
        char names[1024];
        float ms[32];
        uint64_t ns = 0;
        if (hdg_last_stage_ms(index, names, sizeof names, ms, &ns) == HDG_OK) {
          std::printf("device: 0, data batch %llu, queries %llu..%llu stages (ms):", (unsigned long long)b,
                      (unsigned long long)q0, (unsigned long long)(q0 + qn));
          std::string nm(names);
          size_t p = 0;
          for (uint64_t i = 0; i < ns && i < 32; ++i) {
            size_t e = nm.find(',', p);
            std::printf(" %s=%.3f", nm.substr(p, e == std::string::npos ? std::string::npos : e - p).c_str(), ms[i]);
            p = e == std::string::npos ? nm.size() : e + 1;
          }
          std::printf("\n");
        }
      }
      if (n_batches == 1) {
        // results of this block go to the host while the next block computes
        SE_HIP(hipEventRecord(done[bk], comp));
        SE_HIP(hipStreamWaitEvent(copy, done[bk], 0));
        SE_HIP(hipMemcpyAsync(h_dist + q0 * topk, bd + q0 * topk, qn * topk * 4, hipMemcpyDeviceToHost, copy));
        SE_HIP(hipMemcpyAsync(h_ids + q0 * topk, bi + q0 * topk, qn * topk * 8, hipMemcpyDeviceToHost, copy));
      }
    }
    SE_HIP(hipStreamSynchronize(comp));
    query_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
  }
  // ---- 3. merge across data batches on the device ----
  auto t2 = std::chrono::steady_clock::now();
  if (n_batches > 1) {
    SE_HIP(hipMalloc((void**)&dv.mdist, res_n * 4));
    SE_HIP(hipMalloc((void**)&dv.mids, res_n * 8));
    if (hdg_merge_topk(dv.dist, dv.ids, n_batches, db, n_data, 0, nq, topk, dv.mdist, dv.mids, comp) != HDG_OK)
      return die(hdg_last_error());
    SE_HIP(hipMemcpyAsync(h_dist, dv.mdist, res_n * 4, hipMemcpyDeviceToHost, comp));
    SE_HIP(hipMemcpyAsync(h_ids, dv.mids, res_n * 8, hipMemcpyDeviceToHost, comp));
    SE_HIP(hipStreamSynchronize(comp));
  }
  SE_HIP(hipStreamSynchronize(copy));
  query_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t2).count();

  // ---- 4. result files and recall ----
  ::mkdir("res", 0755);
  std::ofstream("res/distances_0", std::ios::binary | std::ios::trunc).write((const char*)h_dist, res_n * 4);
  std::ofstream("res/neighbors_0", std::ios::binary | std::ios::trunc).write((const char*)h_ids, res_n * 8);
  std::printf("build time: %.3f s, query time: %.3f s, QPS: %.1f\n", build_s, query_s, nq / query_s);

  std::ifstream gt(dir + "neighbors", std::ios::binary);
  if (gt) {
    std::vector<int32_t> g(res_n, -1);
    gt.read(reinterpret_cast<char*>(g.data()), g.size() * sizeof(int32_t));
    uint64_t hits = 0, valid = 0;
    for (uint64_t q = 0; q < nq; ++q) {
      std::unordered_set<int64_t> gs, seen;
      for (uint64_t j = 0; j < topk; ++j) {
        int32_t x = g[q * topk + j];
        if (x >= 0 && gs.insert(x).second) ++valid;
      }
      for (uint64_t j = 0; j < topk; ++j) {
        int64_t x = (int64_t)h_ids[q * topk + j];
        if (!seen.insert(x).second) continue;
        hits += gs.count(x);
      }
    }
    std::printf("final recall: %.6f\n", valid ? (double)hits / valid : 0.0);
  }
  for (auto& e : done) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(comp);
  (void)hipStreamDestroy(copy);
  (void)hipHostFree(h_dist);
  (void)hipHostFree(h_ids);
  hdg_destroy(index);
  std::printf("search finish\n");
  return 0;
}
