// Index state checkpoint: one .npy file per array plus meta.json.  The reference keeps no index on
// disk (it rebuilds per data batch, search-engine-auto-batch.cu:265-275); this state is what a
// resumed search and the CPU checker both read.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <sys/stat.h>

#include "hdg_index.hpp"

namespace hdg {

namespace {

template <class T>
const char* npy_descr();
template <>
const char* npy_descr<float>() { return "<f4"; }
template <>
const char* npy_descr<uint32_t>() { return "<u4"; }
template <>
const char* npy_descr<uint64_t>() { return "<u8"; }
template <>
const char* npy_descr<uint8_t>() { return "|u1"; }

template <class T>
void write_npy(const std::string& path, const T* data, const std::vector<uint64_t>& shape) {
  std::string hdr = std::string("{'descr': '") + npy_descr<T>() + "', 'fortran_order': False, 'shape': (";
  uint64_t n = 1;
  for (size_t i = 0; i < shape.size(); ++i) {
    hdr += std::to_string(shape[i]) + (shape.size() == 1 ? "," : (i + 1 < shape.size() ? ", " : ""));
    n *= shape[i];
  }
  hdr += "), }";
  size_t total = 10 + hdr.size() + 1;
  size_t pad = (64 - total % 64) % 64;
  hdr += std::string(pad, ' ') + "\n";
  std::ofstream f(path, std::ios::binary);
  if (!f) fail(kIoError, "cannot write " + path);
  const char magic[] = "\x93NUMPY\x01\x00";
  f.write(magic, 8);
  uint16_t hl = (uint16_t)hdr.size();
  f.write(reinterpret_cast<const char*>(&hl), 2);
  f.write(hdr.data(), hdr.size());
  f.write(reinterpret_cast<const char*>(data), n * sizeof(T));
  if (!f) fail(kIoError, "short write " + path);
}

template <class T>
std::vector<T> read_npy(const std::string& path, std::vector<uint64_t>& shape) {
  std::ifstream f(path, std::ios::binary);
  if (!f) fail(kIoError, "cannot read " + path);
  char magic[8];
  f.read(magic, 8);
  if (std::memcmp(magic, "\x93NUMPY", 6) != 0) fail(kIoError, "not an npy file: " + path);
  uint32_t hl = 0;
  if (magic[6] == 1) {
    uint16_t h16;
    f.read(reinterpret_cast<char*>(&h16), 2);
    hl = h16;
  } else {
    f.read(reinterpret_cast<char*>(&hl), 4);
  }
  std::string hdr(hl, ' ');
  f.read(&hdr[0], hl);
  if (hdr.find(npy_descr<T>()) == std::string::npos) fail(kIoError, "unexpected dtype in " + path);
  if (hdr.find("'fortran_order': False") == std::string::npos) fail(kIoError, "fortran order in " + path);
  size_t a = hdr.find('('), b = hdr.find(')');
  shape.clear();
  std::string dims = hdr.substr(a + 1, b - a - 1);
  std::stringstream ss(dims);
  std::string tok;
  uint64_t n = 1;
  while (std::getline(ss, tok, ',')) {
    size_t p = tok.find_first_not_of(' ');
    if (p == std::string::npos) continue;
    shape.push_back(std::stoull(tok.substr(p)));
    n *= shape.back();
  }
  std::vector<T> v(n);
  f.read(reinterpret_cast<char*>(v.data()), n * sizeof(T));
  if (!f) fail(kIoError, "short read " + path);
  return v;
}

template <class T>
std::vector<T> download(const DevBuf& b, uint64_t n) {
  std::vector<T> v(n);
  if (n) HDG_HIP(hipMemcpy(v.data(), b.data(), n * sizeof(T), hipMemcpyDeviceToHost));
  return v;
}

std::string json_array(const uint64_t* v, int n) {
  std::string s = "[";
  for (int i = 0; i < n; ++i) s += (i ? ", " : "") + std::to_string(v[i]);
  return s + "]";
}
std::string json_array(const int* v, int n) {
  std::string s = "[";
  for (int i = 0; i < n; ++i) s += (i ? ", " : "") + std::to_string(v[i]);
  return s + "]";
}

// tiny reader for the flat meta.json written below: "key": value | [v, v, ...]
std::vector<uint64_t> json_get(const std::string& js, const std::string& key) {
  size_t p = js.find("\"" + key + "\"");
  if (p == std::string::npos) fail(kIoError, "meta.json lacks " + key);
  p = js.find(':', p) + 1;
  while (js[p] == ' ') ++p;
  std::vector<uint64_t> out;
  if (js[p] == '[') {
    size_t e = js.find(']', p);
    std::stringstream ss(js.substr(p + 1, e - p - 1));
    std::string tok;
    while (std::getline(ss, tok, ','))
      if (tok.find_first_not_of(' ') != std::string::npos) out.push_back(std::stoull(tok));
  } else {
    out.push_back(std::stoull(js.substr(p)));
  }
  return out;
}

}  // namespace

void Index::save(const std::string& dir) const {
  if (!st_.built) fail(kBadState, "save before build");
  ::mkdir(dir.c_str(), 0755);
  const IndexState& S = st_;
  const GridShape& g = S.grid;
  HDG_HIP(hipSetDevice(device_));
  HDG_HIP(hipDeviceSynchronize());
  std::ostringstream js;
  js << "{\n  \"format\": 1,\n  \"N\": " << S.N << ",\n  \"d\": " << S.d << ",\n  \"C\": " << S.C
     << ",\n  \"CS\": " << S.CS << ",\n  \"M\": " << S.M << ",\n  \"K\": " << S.K << ",\n  \"pq_len\": " << S.pq_len
     << ",\n  \"l\": " << S.l << ",\n  \"dims\": " << json_array(g.dims, g.l)
     << ",\n  \"span\": " << json_array(g.span, g.l) << ",\n  \"left\": " << json_array(g.left, g.l)
     << ",\n  \"shuffle\": " << json_array(g.shuffle, g.l) << ",\n  \"max_bits\": " << g.max_bits
     << ",\n  \"valid_map\": " << g.valid_map << ",\n  \"tot_grids\": " << g.tot_grids
     << ",\n  \"n_valid\": " << g.n_valid << "\n}\n";
  std::ofstream(dir + "/meta.json") << js.str();
  std::vector<uint32_t> seg32(S.h_seg_off.begin(), S.h_seg_off.end());
  write_npy(dir + "/zid.npy", S.h_zid.data(), {S.N});
  write_npy(dir + "/zcode.npy", S.h_zcode.data(), {S.N});
  write_npy(dir + "/seg_off.npy", S.h_seg_off.data(), {S.CS + 1});
  write_npy(dir + "/sub_off.npy", S.h_sub_off.data(), {S.C + 1});
  write_npy(dir + "/centers.npy", download<float>(S.centers, S.C * S.d).data(), {S.C, S.d});
  write_npy(dir + "/sub_centers.npy", download<float>(S.sub_centers, S.CS * S.d).data(), {S.CS, S.d});
  write_npy(dir + "/pq_centers.npy", download<float>(S.pq_centers, S.M * S.K * S.pq_len).data(),
            {S.M, S.K, S.pq_len});
  write_npy(dir + "/codes.npy", download<uint8_t>(S.codes, S.N * S.M).data(), {S.N, S.M});
  write_npy(dir + "/cells.npy", download<uint32_t>(S.cells, (uint64_t)g.n_valid * S.N).data(),
            {(uint64_t)g.n_valid, S.N});
  write_npy(dir + "/vec.npy", download<float>(S.vec, S.N * S.d).data(), {S.N, S.d});
  write_npy(dir + "/lab.npy", download<float>(S.lab, S.N * (uint64_t)S.l).data(), {S.N, (uint64_t)S.l});
}

void Index::load(const std::string& dir, hipStream_t s) {
  HDG_HIP(hipSetDevice(device_));
  drop_graphs();
  std::ifstream mf(dir + "/meta.json");
  if (!mf) fail(kIoError, "cannot read " + dir + "/meta.json");
  std::stringstream ss;
  ss << mf.rdbuf();
  const std::string js = ss.str();
  IndexState S;
  S.N = json_get(js, "N")[0];
  S.d = json_get(js, "d")[0];
  S.C = json_get(js, "C")[0];
  S.CS = json_get(js, "CS")[0];
  S.M = json_get(js, "M")[0];
  S.K = json_get(js, "K")[0];
  S.pq_len = json_get(js, "pq_len")[0];
  S.l = (int)json_get(js, "l")[0];
  if (S.l != fc_.l) fail(kBadState, "saved index has a different number of label dims than filter.conf");
  GridShape& g = S.grid;
  g.l = S.l;
  auto dims = json_get(js, "dims"), span = json_get(js, "span"), left = json_get(js, "left"),
       shuf = json_get(js, "shuffle");
  for (int i = 0; i < g.l; ++i) {
    g.dims[i] = (int)dims[i];
    g.span[i] = span[i];
    g.left[i] = left[i];
    g.shuffle[i] = (int)shuf[i];
  }
  g.max_bits = (int)json_get(js, "max_bits")[0];
  g.valid_map = (int)json_get(js, "valid_map")[0];
  g.tot_grids = json_get(js, "tot_grids")[0];
  g.n_valid = 0;
  for (int i = 0; i < g.l; ++i)
    if ((g.valid_map >> i) & 1) g.valid_dims[g.n_valid++] = i;
  std::vector<uint64_t> sh;
  S.h_zid = read_npy<uint32_t>(dir + "/zid.npy", sh);
  S.h_zcode = read_npy<uint64_t>(dir + "/zcode.npy", sh);
  S.h_seg_off = read_npy<uint64_t>(dir + "/seg_off.npy", sh);
  S.h_sub_off = read_npy<uint32_t>(dir + "/sub_off.npy", sh);
  if (S.h_seg_off.size() != S.CS + 1 || S.h_sub_off.size() != S.C + 1 || S.h_zid.size() != S.N)
    fail(kIoError, "inconsistent index arrays in " + dir);
  for (uint64_t c = 0; c < S.C; ++c)
    S.max_subs_per_cluster = std::max<int>(S.max_subs_per_cluster, (int)(S.h_sub_off[c + 1] - S.h_sub_off[c]));
  for (uint64_t i = 0; i < S.CS; ++i)
    S.max_seg_len = std::max<uint32_t>(S.max_seg_len, (uint32_t)(S.h_seg_off[i + 1] - S.h_seg_off[i]));
  upload(S.zid, S.h_zid, s);
  upload(S.zcode, S.h_zcode, s);
  upload(S.sub_off, S.h_sub_off, s);
  {
    std::vector<uint32_t> so(S.h_seg_off.begin(), S.h_seg_off.end());
    upload(S.seg_off, so, s);
    std::vector<uint32_t> cos(S.CS);
    for (uint64_t c = 0; c < S.C; ++c)
      for (uint32_t p = S.h_sub_off[c]; p < S.h_sub_off[c + 1]; ++p) cos[p] = (uint32_t)c;
    upload(S.cluster_of_sub, cos, s);
  }
  {
    std::vector<float> centers = read_npy<float>(dir + "/centers.npy", sh);
    std::vector<float> subc = read_npy<float>(dir + "/sub_centers.npy", sh);
    upload(S.centers, centers, s);
    upload(S.sub_centers, subc, s);
    prepare_mixed_centers(S, centers, subc, s);
  }
  upload(S.pq_centers, read_npy<float>(dir + "/pq_centers.npy", sh), s);
  upload(S.codes, read_npy<uint8_t>(dir + "/codes.npy", sh), s);
  {
    std::vector<uint32_t> cells = read_npy<uint32_t>(dir + "/cells.npy", sh);
    if (cells.size() != (uint64_t)g.n_valid * S.N) fail(kIoError, "inconsistent cells array in " + dir);
    upload(S.cells, cells, s);
    build_bucket_grid(S, cells, s);
  }
  upload(S.vec, read_npy<float>(dir + "/vec.npy", sh), s);
  {
    std::vector<float> lab = read_npy<float>(dir + "/lab.npy", sh);
    S.labels_integral = std::all_of(lab.begin(), lab.end(), [](float v) { return v == std::floor(v); });
    upload(S.lab, lab, s);
  }
  S.built = true;
  st_ = std::move(S);
}

}  // namespace hdg
