// Index state checkpoint: one .npy file per array plus meta.json.  The reference keeps no index on
// disk (it rebuilds per data batch, search-engine-auto-batch.cu:265-275); this state is what a
// resumed search and the CPU checker both read.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
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
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) fail(kIoError, "cannot read " + path);
  const uint64_t fsize = (uint64_t)f.tellg();
  f.seekg(0);
  char magic[8];
  f.read(magic, 8);
  if (!f || std::memcmp(magic, "\x93NUMPY", 6) != 0) fail(kIoError, "not an npy file: " + path);
  uint32_t hl = 0;
  if (magic[6] == 1) {
    uint16_t h16 = 0;
    f.read(reinterpret_cast<char*>(&h16), 2);
    hl = h16;
  } else {
    f.read(reinterpret_cast<char*>(&hl), 4);
  }
  if (!f || hl > fsize) fail(kIoError, "bad npy header in " + path);
  std::string hdr(hl, ' ');
  f.read(&hdr[0], hl);
  if (!f) fail(kIoError, "short npy header in " + path);
  if (hdr.find(npy_descr<T>()) == std::string::npos) fail(kIoError, "unexpected dtype in " + path);
  if (hdr.find("'fortran_order': False") == std::string::npos) fail(kIoError, "fortran order in " + path);
  const size_t a = hdr.find('('), b = hdr.find(')');
  if (a == std::string::npos || b == std::string::npos || b < a) fail(kIoError, "bad npy shape in " + path);
  shape.clear();
  std::string dims = hdr.substr(a + 1, b - a - 1);
  std::stringstream ss(dims);
  std::string tok;
  uint64_t n = 1;
  while (std::getline(ss, tok, ',')) {
    size_t p = tok.find_first_not_of(' ');
    if (p == std::string::npos) continue;
    if (tok.find_first_not_of("0123456789 ", p) != std::string::npos) fail(kIoError, "bad npy shape in " + path);
    shape.push_back(std::stoull(tok.substr(p)));
    if (shape.back() && n > ~0ull / shape.back()) fail(kIoError, "npy shape overflows in " + path);
    n *= shape.back();
  }
  // the data must be exactly what the shape says: a truncated or padded file is refused before allocating
  const uint64_t data_off = (uint64_t)f.tellg();
  if (n > ~0ull / sizeof(T) || fsize - data_off != n * sizeof(T))
    fail(kIoError, "npy data size does not match its shape in " + path);
  std::vector<T> v(n);
  f.read(reinterpret_cast<char*>(v.data()), n * sizeof(T));
  if (!f) fail(kIoError, "short read " + path);
  return v;
}

// Throws kIoError unless an array read from `name` has exactly the expected shape.
void expect_shape(const std::vector<uint64_t>& got, const std::vector<uint64_t>& want, const std::string& name) {
  if (got != want) fail(kIoError, "array " + name + " of the saved index has the wrong shape");
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
  p = js.find(':', p);
  if (p == std::string::npos) fail(kIoError, "meta.json: no value for " + key);
  ++p;
  while (p < js.size() && js[p] == ' ') ++p;
  std::vector<uint64_t> out;
  auto num = [&](const std::string& t) -> uint64_t {
    const size_t q = t.find_first_not_of(" \n\r\t");
    if (q == std::string::npos || !std::isdigit((unsigned char)t[q])) fail(kIoError, "meta.json: bad value of " + key);
    return std::stoull(t.substr(q));
  };
  if (p < js.size() && js[p] == '[') {
    size_t e = js.find(']', p);
    if (e == std::string::npos) fail(kIoError, "meta.json: unterminated array " + key);
    std::stringstream ss(js.substr(p + 1, e - p - 1));
    std::string tok;
    while (std::getline(ss, tok, ','))
      if (tok.find_first_not_of(' ') != std::string::npos) out.push_back(num(tok));
  } else {
    out.push_back(num(js.substr(p, 32)));
  }
  return out;
}

uint64_t json_scalar(const std::string& js, const std::string& key) {
  auto v = json_get(js, key);
  if (v.size() != 1) fail(kIoError, "meta.json: " + key + " is not a scalar");
  return v[0];
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

// Reads a saved index state into host memory.  Every array's shape is checked against meta.json and every
// value that is used as an index or offset is range-checked, so a truncated or inconsistent checkpoint fails
// with kIoError here instead of reaching the query kernels.  Touches no device.
HostCheckpoint read_checkpoint(const std::string& dir, int l_expected) {
  HostCheckpoint H;
  IndexState& S = H.S;
  std::ifstream mf(dir + "/meta.json");
  if (!mf) fail(kIoError, "cannot read " + dir + "/meta.json");
  std::stringstream ss;
  ss << mf.rdbuf();
  const std::string js = ss.str();
  S.N = json_scalar(js, "N");
  S.d = json_scalar(js, "d");
  S.C = json_scalar(js, "C");
  S.CS = json_scalar(js, "CS");
  S.M = json_scalar(js, "M");
  S.K = json_scalar(js, "K");
  S.pq_len = json_scalar(js, "pq_len");
  const uint64_t l64 = json_scalar(js, "l");
  auto bad = [&](const std::string& what) { fail(kIoError, "saved index in " + dir + ": " + what); };
  if (l64 < 1 || l64 > (uint64_t)kMaxDims) bad("label dims outside [1, 8]");
  S.l = (int)l64;
  if (S.l != l_expected) fail(kBadState, "saved index has a different number of label dims than filter.conf");
  if (S.N == 0 || S.N >= (1ull << 32)) bad("row count outside [1, 2^32)");
  if (S.d == 0 || S.C == 0 || S.CS < S.C || S.CS >= (1ull << 32) || S.M == 0 || S.M > S.d) bad("bad sizes");
  if (S.K == 0 || S.K > 256) bad("PQ codewords outside [1, 256]");
  if (S.pq_len != (S.d + S.M - 1) / S.M) bad("pq_len differs from ceil(d / M)");
  GridShape& g = S.grid;
  g.l = S.l;
  auto dims = json_get(js, "dims"), span = json_get(js, "span"), left = json_get(js, "left"),
       shuf = json_get(js, "shuffle");
  if (dims.size() != l64 || span.size() != l64 || left.size() != l64 || shuf.size() != l64)
    bad("grid arrays of meta.json do not have l entries");
  uint32_t seen = 0;
  for (int i = 0; i < g.l; ++i) {
    if (dims[i] == 0 || dims[i] > 0x7fffffffull || shuf[i] >= l64 || ((seen >> shuf[i]) & 1))
      bad("bad grid shape");
    seen |= 1u << shuf[i];
    g.dims[i] = (int)dims[i];
    g.span[i] = span[i];
    g.left[i] = left[i];
    g.shuffle[i] = (int)shuf[i];
  }
  const uint64_t max_bits = json_scalar(js, "max_bits"), valid_map = json_scalar(js, "valid_map");
  if (max_bits < 1 || max_bits > 32 || valid_map == 0 || valid_map >= (1ull << g.l)) bad("bad Morton configuration");
  g.max_bits = (int)max_bits;
  g.valid_map = (int)valid_map;
  g.tot_grids = json_scalar(js, "tot_grids");
  g.n_valid = 0;
  for (int i = 0; i < g.l; ++i)
    if ((g.valid_map >> i) & 1) g.valid_dims[g.n_valid++] = i;
  for (int j = 0; j < g.n_valid; ++j) {
    if (g.span[g.valid_dims[j]] == 0) bad("a valid dim has a zero label span");
    if (((uint64_t)g.dims[g.valid_dims[j]] - 1) >> max_bits) bad("a valid dim has more cells than its Morton bits hold");
  }
  if ((uint64_t)g.n_valid * max_bits > 64) bad("Morton codes wider than 64 bits");
  std::vector<uint64_t> sh;
  S.h_zid = read_npy<uint32_t>(dir + "/zid.npy", sh);
  expect_shape(sh, {S.N}, "zid");
  S.h_zcode = read_npy<uint64_t>(dir + "/zcode.npy", sh);
  expect_shape(sh, {S.N}, "zcode");
  S.h_seg_off = read_npy<uint64_t>(dir + "/seg_off.npy", sh);
  expect_shape(sh, {S.CS + 1}, "seg_off");
  S.h_sub_off = read_npy<uint32_t>(dir + "/sub_off.npy", sh);
  expect_shape(sh, {S.C + 1}, "sub_off");
  if (S.h_seg_off[0] != 0 || S.h_seg_off[S.CS] != S.N) bad("seg_off does not span [0, N]");
  if (S.h_sub_off[0] != 0 || S.h_sub_off[S.C] != S.CS) bad("sub_off does not span [0, CS]");
  for (uint64_t c = 0; c < S.C; ++c) {
    if (S.h_sub_off[c + 1] < S.h_sub_off[c]) bad("sub_off is not monotone");
    S.max_subs_per_cluster = std::max<int>(S.max_subs_per_cluster, (int)(S.h_sub_off[c + 1] - S.h_sub_off[c]));
  }
  for (uint64_t i = 0; i < S.CS; ++i) {
    if (S.h_seg_off[i + 1] < S.h_seg_off[i]) bad("seg_off is not monotone");
    S.max_seg_len = std::max<uint32_t>(S.max_seg_len, (uint32_t)(S.h_seg_off[i + 1] - S.h_seg_off[i]));
    for (uint64_t p = S.h_seg_off[i] + 1; p < S.h_seg_off[i + 1]; ++p)
      if (S.h_zcode[p] < S.h_zcode[p - 1]) bad("Morton codes are not ascending inside a sub-cluster");
  }
  {
    std::vector<uint8_t> hit(S.N, 0);
    for (uint32_t z : S.h_zid) {
      if (z >= S.N || hit[z]) bad("zid is not a permutation of the rows");
      hit[z] = 1;
    }
  }
  H.centers = read_npy<float>(dir + "/centers.npy", sh);
  expect_shape(sh, {S.C, S.d}, "centers");
  H.subc = read_npy<float>(dir + "/sub_centers.npy", sh);
  expect_shape(sh, {S.CS, S.d}, "sub_centers");
  H.pqc = read_npy<float>(dir + "/pq_centers.npy", sh);
  expect_shape(sh, {S.M, S.K, S.pq_len}, "pq_centers");
  H.codes = read_npy<uint8_t>(dir + "/codes.npy", sh);
  expect_shape(sh, {S.N, S.M}, "codes");
  for (uint8_t c : H.codes)
    if (c >= S.K) bad("a PQ code is not below the codeword count");
  H.cells = read_npy<uint32_t>(dir + "/cells.npy", sh);
  expect_shape(sh, {(uint64_t)g.n_valid, S.N}, "cells");
  for (int j = 0; j < g.n_valid; ++j) {
    const uint32_t lim = (uint32_t)g.dims[g.valid_dims[j]];
    for (uint64_t p = 0; p < S.N; ++p)
      if (H.cells[(uint64_t)j * S.N + p] >= lim) bad("a grid cell lies outside its dim");
  }
  H.vec = read_npy<float>(dir + "/vec.npy", sh);
  expect_shape(sh, {S.N, S.d}, "vec");
  H.lab = read_npy<float>(dir + "/lab.npy", sh);
  expect_shape(sh, {S.N, (uint64_t)S.l}, "lab");
  return H;
}

void Index::load(const std::string& dir, hipStream_t s) {
  HDG_HIP(hipSetDevice(device_));
  HostCheckpoint H = read_checkpoint(dir, fc_.l);
  IndexState& S = H.S;
  // everything is consistent: only now is the current index replaced
  drop_graphs();
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
  upload(S.centers, H.centers, s);
  upload(S.sub_centers, H.subc, s);
  prepare_mixed_centers(S, H.centers, H.subc, s);
  upload(S.pq_centers, H.pqc, s);
  upload(S.codes, H.codes, s);
  prepare_blocked_codes(S, s);
  upload(S.cells, H.cells, s);
  build_bucket_grid(S, s);
  upload(S.vec, H.vec, s);
  upload(S.lab, H.lab, s);
  prepare_label_words(S, H.lab, s);
  S.built = true;
  st_ = std::move(S);
}

}  // namespace hdg
