// Super-Grid query path on the GPU: Super-Grid-IVF_PQ.cu:985-1248 (index::query) and the kernels it
// reaches, redesigned for gfx950.  Stage map (reference row -> kernel here):
//   A1 filter recovery + A2 grid boxes + A7 Morton corners  -> prep_kernel
//   A3 in-box counts per (query, sub-cluster)               -> count_kernel (streams packed cells)
//   A4/A5 non-zero clusters + exact coarse top-n_list       -> cluster_nnz_kernel, coarse_select_kernel
//   A6 non-zero sub-clusters + exact top-sub_lists          -> sub_nnz_kernel, sub_select_kernel
//   A8/A9 Morton ranges, candidate counts, beta sums        -> ranges_kernel
//   A10 beta / n_candies, buffer plan                       -> plan_kernel + scan
//   A11 PQ ADC scan + top-k' (only where C_q > k')          -> adc_scan_kernel, adc_select_kernel
//   A12 label post-filter (l > 1)                           -> inside keep_all / adc_select
//   A13 exact FP32 re-rank + final top-k                    -> final_kernel
// Arithmetic that decides integer results (cells, Morton codes, counts) and the FP32 distances follow
// the reference's operation order; the file is compiled with -ffp-contract=off and every FMA that the
// reference's compiler contracts is an explicit fmaf().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cstring>

#include <hip/hip_fp16.h>

#include "hdg_device.hpp"
#include "hdg_fast.hpp"
#include "hdg_index.hpp"

namespace hdg {

using dev::kWG;

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr float kExactPad = 1e6f;               // refine.cu:53
constexpr uint64_t kIdPad = 4294967296ull;      // mat_operators.cuh:185
constexpr uint64_t kU64Max = ~0ull;             // filter.cu:171
constexpr int kCountSegTile = 32;               // segments per LDS flush in count_kernel
constexpr int kCountSegChunk = 64;              // segments per count_kernel block
constexpr uint32_t kAdcChunk = 8192;            // candidates per adc_scan block

struct Dims {
  int l, n_valid, max_bits, valid_map;
  int dims[kMaxDims];
  int shuffle[kMaxDims];
  uint64_t span[kMaxDims];
  uint64_t left[kMaxDims];
  int valid_dims[kMaxDims];
};

Dims make_dims(const GridShape& g) {
  Dims D;
  D.l = g.l;
  D.n_valid = g.n_valid;
  D.max_bits = g.max_bits;
  D.valid_map = g.valid_map;
  for (int i = 0; i < kMaxDims; ++i) {
    D.dims[i] = g.dims[i];
    D.shuffle[i] = g.shuffle[i];
    D.span[i] = g.span[i];
    D.left[i] = g.left[i];
    D.valid_dims[i] = g.valid_dims[i];
  }
  return D;
}

__host__ __device__ inline uint64_t cdiv(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// ------------------------------------------------------------------------------------------------
// A1 + A2 + A7.  One thread per query.  Query qg belongs to group g = qg / G; the recovery kernel of
// the reference indexes the group's [G][qcols] view as a flat array (preprocessing.cu:10-30).
__global__ void prep_kernel(const float* __restrict__ qlabels, uint64_t qcols, uint64_t q0, uint64_t qn,
                            uint64_t G, Dims D, const int* __restrict__ ftype,
                            const float* __restrict__ shift, const float* __restrict__ imap,
                            uint64_t imap_len, const int* __restrict__ divv, float* __restrict__ ranges,
                            int* __restrict__ boxes, uint32_t* __restrict__ box_lo,
                            uint32_t* __restrict__ box_w, uint64_t* __restrict__ zlo,
                            uint64_t* __restrict__ zhi, uint64_t* __restrict__ zlo_cnt,
                            uint32_t* __restrict__ err) {
  uint64_t qi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (qi >= qn) return;
  const uint64_t qg = q0 + qi;
  const uint64_t b0 = (qg / G) * G, t = qg - b0;
  const float* base = qlabels + b0 * qcols;
  const int l = D.l;
  float rg[2 * kMaxDims];
  for (int i = 0; i < l; ++i) {
    const uint64_t idx = t * (uint64_t)l + (uint64_t)i;
    float lo = 0.0f, hi = 0.0f;
    switch (ftype[i]) {
      case 0: {
        // CUDA max(float, double) promotes to double; the result is rounded to f32 on store
        double s0 = fmax((double)shift[2 * i], 1e-7), s1 = fmax((double)shift[2 * i + 1], 1e-7);
        lo = (float)((double)base[idx] - s0);
        hi = (float)((double)base[idx] + s1);
        break;
      }
      case 1: {
        int v = (int)(base[idx] - 1.0f);
        if (v < 0 || (uint64_t)(2 * (int64_t)v + 1) >= imap_len) {
          atomicOr(err, 1u);
          v = 0;
        }
        lo = imap[2 * v];
        hi = imap[2 * v + 1];
        break;
      }
      case 2: {
        int v = (int)base[idx];
        lo = (float)(v / divv[i]);
        hi = FLT_MAX;
        break;
      }
      default:  // 3: explicit per-query range
        lo = base[2 * idx];
        hi = base[2 * idx + 1];
        break;
    }
    rg[2 * i] = lo;
    rg[2 * i + 1] = hi;
    ranges[qi * 2 * l + 2 * i] = lo;
    ranges[qi * 2 * l + 2 * i + 1] = hi;
  }
  // A2: Super-Grid-IVF_PQ.cu:947-960
  int lo_c[kMaxDims], hi_c[kMaxDims];
  for (int i = 0; i < l; ++i) {
    const int d = D.shuffle[i];
    uint64_t a = f2u_sat(rg[2 * d]), b = f2u_sat(rg[2 * d + 1]);
    float fl = (float)(a > D.left[i] ? a : D.left[i]);
    uint64_t top = D.left[i] + D.span[i];
    float fr = (float)(b < top ? b : top);
    int gl = grid_cell(fl, D.left[i], D.span[i], D.dims[i]);
    int gr = grid_cell(fr, D.left[i], D.span[i], D.dims[i]);
    int h = gr < D.dims[i] - 1 ? gr : D.dims[i] - 1;
    int lw = h < gl ? h : gl;
    hi_c[i] = h;
    lo_c[i] = lw;
    boxes[qi * 2 * l + 2 * i] = lw;
    boxes[qi * 2 * l + 2 * i + 1] = h;
  }
  for (int j = 0; j < D.n_valid; ++j) {
    int i = D.valid_dims[j];
    box_lo[(uint64_t)j * qn + qi] = (uint32_t)lo_c[i];
    box_w[(uint64_t)j * qn + qi] = (uint32_t)hi_c[i] - (uint32_t)lo_c[i];
  }
  zlo[qi] = morton_code(lo_c, l, D.max_bits, D.valid_map);
  zhi[qi] = morton_code(hi_c, l, D.max_bits, D.valid_map);
  // Morton range used by the count: a lower corner that wrapped negative (a truncated u64 cell id)
  // still admits every cell >= 0, so its code is taken at 0.
  for (int i = 0; i < l; ++i) lo_c[i] = lo_c[i] < 0 ? 0 : lo_c[i];
  zlo_cnt[qi] = morton_code(lo_c, l, D.max_bits, D.valid_map);
}

// ------------------------------------------------------------------------------------------------
// A3: counts[q][s] = #points of sub-cluster s whose cells lie in q's box on every valid dim.  This is
// the integer the reference obtains by 2^l-corner inclusion-exclusion over compressed prefix sums
// (Super-Grid-IVF_PQ.cu:648-790).  Here it is counted directly: a point inside the box has a Morton
// code in [z(lo), z(hi)] (the bit-interleave is monotone in every coordinate), and a sub-cluster's
// points are sorted by Morton code, so only that code range of each segment is box-tested.
// Block = 256 lanes = 256 queries; blockIdx.y = chunk of kCountSegChunk segments.  Each segment is
// staged in LDS in pieces of kCountPiece points (codes + cells); every lane binary-searches its own
// code range in LDS and tests the points inside it.
constexpr int kCountPiece = 1024;

template <int NV>
__global__ __launch_bounds__(256) void count_kernel(const uint32_t* __restrict__ cells, uint64_t N,
                                                    const uint64_t* __restrict__ zcode,
                                                    const uint32_t* __restrict__ seg_off, uint64_t CS,
                                                    const uint32_t* __restrict__ box_lo,
                                                    const uint32_t* __restrict__ box_w,
                                                    const uint64_t* __restrict__ zlo,
                                                    const uint64_t* __restrict__ zhi, uint64_t qn,
                                                    uint32_t* __restrict__ counts) {
  __shared__ uint64_t zs[kCountPiece];
  __shared__ uint32_t cs[NV][kCountPiece];
  const uint64_t q = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  const bool live = q < qn;
  uint32_t lo[NV], w[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    lo[j] = live ? box_lo[(uint64_t)j * qn + q] : 1u;
    w[j] = live ? box_w[(uint64_t)j * qn + q] : 0u;
  }
  const uint64_t za = live ? zlo[q] : ~0ull, zb = live ? zhi[q] : 0ull;
  const uint64_t s_begin = (uint64_t)blockIdx.y * kCountSegChunk;
  const uint64_t s_end = s_begin + kCountSegChunk < CS ? s_begin + kCountSegChunk : CS;
  for (uint64_t s = s_begin; s < s_end; ++s) {
    const uint32_t p0 = seg_off[s], p1 = seg_off[s + 1];
    uint32_t cnt = 0;
    for (uint32_t base = p0; base < p1; base += kCountPiece) {
      const uint32_t m = p1 - base < (uint32_t)kCountPiece ? p1 - base : (uint32_t)kCountPiece;
      __syncthreads();
      for (uint32_t i = threadIdx.x; i < m; i += 256) {
        zs[i] = zcode[base + i];
#pragma unroll
        for (int j = 0; j < NV; ++j) cs[j][i] = cells[(uint64_t)j * N + base + i];
      }
      __syncthreads();
      if (za > zs[m - 1] || zb < zs[0]) continue;
      uint32_t a = 0, b = m;  // first code >= za
      while (a < b) {
        uint32_t mid = (a + b) >> 1;
        if (zs[mid] < za) a = mid + 1; else b = mid;
      }
      uint32_t c = a, e = m;  // first code > zb
      while (c < e) {
        uint32_t mid = (c + e) >> 1;
        if (zs[mid] <= zb) c = mid + 1; else e = mid;
      }
      for (uint32_t i = a; i < c; ++i) {
        bool in = true;
#pragma unroll
        for (int j = 0; j < NV; ++j) in = in && (cs[j][i] - lo[j] <= w[j]);
        cnt += in ? 1u : 0u;
      }
    }
    if (live) counts[q * CS + s] = cnt;
  }
}

// ------------------------------------------------------------------------------------------------
// Exact FP32 L2 of the workgroup's query against up to 256 rows (one per thread), the reference's
// order for d <= 512 (refine.cu:36-75): a sequential f32 sum over dims of rounded squares.
// rows[] in LDS holds row indices (kNone = no row).  stage: 256 x 33 floats of LDS.
__device__ float exact_tile(const float* __restrict__ X, uint64_t d, const uint32_t* rows,
                            const float* qv, float* stage) {
  const uint32_t my = rows[threadIdx.x];
  float acc = 0.0f;
  for (uint64_t t0 = 0; t0 < d; t0 += 32) {
    const uint32_t w = (uint32_t)(d - t0 < 32 ? d - t0 : 32);
    if ((d & 3) == 0) {
      // 8 lanes per row: each wave load instruction reads 8 whole 128-byte row chunks
#pragma unroll
      for (int it = 0; it < 8; ++it) {
        const uint32_t e = threadIdx.x + it * kWG;
        const uint32_t r = e >> 3, c4 = (e & 7) * 4;
        const uint32_t row = rows[r];
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (row != kNone && c4 < w) v = *reinterpret_cast<const float4*>(X + (uint64_t)row * d + t0 + c4);
        float* st = stage + r * 33 + c4;
        st[0] = v.x;
        st[1] = v.y;
        st[2] = v.z;
        st[3] = v.w;
      }
    } else {
      for (uint32_t e = threadIdx.x; e < 256u * 32u; e += kWG) {
        const uint32_t r = e >> 5, c = e & 31;
        const uint32_t row = rows[r];
        stage[r * 33 + c] = (row != kNone && c < w) ? X[(uint64_t)row * d + t0 + c] : 0.0f;
      }
    }
    __syncthreads();
    if (my != kNone) {
      for (uint32_t c = 0; c < w; ++c) {
        float diff = qv[t0 + c] - stage[threadIdx.x * 33 + c];
        float sq = diff * diff;
        acc = acc + sq;
      }
    }
    __syncthreads();
  }
  return acc;
}

// Sort keys[0..n) (global scratch) ascending into `sorted` (LDS when n <= kSortCap) or, for larger
// n, select the k smallest and sort those.  Returns how many sorted keys are available (min(n,k)).
__device__ uint32_t select_sorted(uint64_t* __restrict__ g, uint32_t n, uint32_t k, uint64_t* lds,
                                  uint32_t* hist, uint64_t* sh) {
  uint32_t m = n < k ? n : k;
  if (n <= (uint32_t)dev::kSortCap) {
    for (uint32_t i = threadIdx.x; i < n; i += kWG) lds[i] = g[i];
    __syncthreads();
    dev::wg_sort_lds(lds, n);
    return m;
  }
  if (m > (uint32_t)dev::kSortCap) m = dev::kSortCap;  // callers never ask for more
  dev::wg_radix_select(g, n, m, lds, hist, sh);
  dev::wg_sort_lds(lds, m);
  return m;
}

// ------------------------------------------------------------------------------------------------
// A4: non-zero clusters of each query, ascending (filter.cu:242-260 over get_cluster_sum_kernel's
// int sums, Super-Grid-IVF_PQ.cu:767-790).  One workgroup per query.
__global__ __launch_bounds__(256) void cluster_nnz_kernel(const uint32_t* __restrict__ counts, uint64_t CS,
                                                          const uint32_t* __restrict__ sub_off, uint32_t C,
                                                          uint64_t G, uint64_t q0,
                                                          uint32_t* __restrict__ list_c,
                                                          uint32_t* __restrict__ nnz_c,
                                                          unsigned long long* __restrict__ gstat) {
  __shared__ uint32_t tmp[8];
  const uint64_t qi = blockIdx.x;
  const uint32_t per = (C + kWG - 1) / kWG;
  const uint32_t c0 = threadIdx.x * per, c1 = min(C, c0 + per);
  const uint32_t* row = counts + qi * CS;
  uint32_t mine = 0;
  for (uint32_t c = c0; c < c1; ++c) {
    int sum = 0;
    for (uint32_t s = sub_off[c]; s < sub_off[c + 1]; ++s) sum += (int)row[s];
    mine += sum != 0;
  }
  uint32_t total;
  uint32_t pos = dev::wg_exclusive_scan(mine, tmp, total);
  for (uint32_t c = c0; c < c1; ++c) {
    int sum = 0;
    for (uint32_t s = sub_off[c]; s < sub_off[c + 1]; ++s) sum += (int)row[s];
    if (sum != 0) list_c[qi * C + pos++] = c;
  }
  if (threadIdx.x == 0) {
    nnz_c[qi] = total;
    atomicMax(&gstat[((q0 + qi) / G) * 8 + 0], (unsigned long long)total);
  }
}

// Number of real items kept by a reference top-k over `real` sorted items followed by `pads` pad
// columns of value 1e6 (pad columns sort after every real column of equal value).
__device__ __forceinline__ uint32_t kept_with_pads(const uint64_t* sorted, uint32_t m, uint32_t k,
                                                   uint64_t pads) {
  // count real items with value <= 1e6 among the sorted prefix
  uint32_t n_le = 0;
  while (n_le < m && dev::key_value(sorted[n_le]) <= kExactPad) ++n_le;
  uint32_t keep = k < n_le ? k : n_le;
  if (k > n_le + pads) keep = n_le + (uint32_t)min<uint64_t>(k - n_le - pads, m - n_le);
  return keep;
}

// A5: exact distances to the non-zero clusters, sorted top kc = min(n_list, group max nnz)
// (Super-Grid-IVF_PQ.cu:1027-1042, refine.cu:78-119 with is_sort = true).
__global__ __launch_bounds__(256) void coarse_select_kernel(
    const float* __restrict__ queries, uint64_t d, const float* __restrict__ centers, uint32_t C,
    uint32_t n_list, uint64_t G, uint64_t q0, const uint32_t* __restrict__ list_c,
    const uint32_t* __restrict__ nnz_c, const unsigned long long* __restrict__ gstat,
    uint64_t* __restrict__ keys, uint32_t* __restrict__ chosen_c, uint32_t* __restrict__ n_chosen_c) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint64_t* sorted = reinterpret_cast<uint64_t*>(smem);               // kSortCap keys
  float* stage = reinterpret_cast<float*>(smem);                      // aliases `sorted` (phase 1)
  __shared__ float qv[512];
  __shared__ uint32_t rows[256];
  __shared__ uint32_t hist[2048];
  __shared__ uint64_t sh[8];
  const uint64_t qi = blockIdx.x;
  const float* q = queries + (q0 + qi) * d;
  for (uint32_t t = threadIdx.x; t < d; t += kWG) qv[t] = q[t];
  const uint32_t n = nnz_c[qi];
  uint64_t* gk = keys + qi * C;
  __syncthreads();
  for (uint32_t b = 0; b < n; b += kWG) {
    const uint32_t j = b + threadIdx.x;
    rows[threadIdx.x] = j < n ? list_c[qi * C + j] : kNone;
    __syncthreads();
    float dist = exact_tile(centers, d, rows, qv, stage);
    if (j < n) gk[j] = dev::make_key(dist, j);
    __syncthreads();
  }
  __syncthreads();
  const uint64_t gmax = gstat[((q0 + qi) / G) * 8 + 0];
  const uint32_t kc = (uint32_t)min<uint64_t>(n_list, gmax);
  const uint32_t m = select_sorted(gk, n, kc, sorted, hist, sh);
  __syncthreads();
  const uint32_t keep = kept_with_pads(sorted, m, kc, (uint64_t)C - n);
  for (uint32_t j = threadIdx.x; j < keep; j += kWG)
    chosen_c[qi * n_list + j] = list_c[qi * C + dev::key_tie(sorted[j])];
  if (threadIdx.x == 0) n_chosen_c[qi] = keep;
}

// ------------------------------------------------------------------------------------------------
// A6 part 1: expand the chosen clusters (in their sorted order) into sub-cluster table positions
// (build_offset_table, mat_operators.cuh:692-794) and keep the non-zero ones in table order
// (collect_segments_data + filter_none_zero_data, Super-Grid-IVF_PQ.cu:866-883).
__global__ __launch_bounds__(256) void sub_nnz_kernel(const uint32_t* __restrict__ counts, uint64_t CS,
                                                      const uint32_t* __restrict__ sub_off,
                                                      uint32_t n_list, uint64_t G, uint64_t q0,
                                                      const uint32_t* __restrict__ chosen_c,
                                                      const uint32_t* __restrict__ n_chosen_c,
                                                      uint32_t cap, uint32_t* __restrict__ list_s,
                                                      uint32_t* __restrict__ nnz_s,
                                                      unsigned long long* __restrict__ gstat) {
  __shared__ uint32_t pre[1025];
  __shared__ uint32_t tmp[8];
  __shared__ uint32_t base_sh;
  const uint64_t qi = blockIdx.x;
  const uint32_t nc = n_chosen_c[qi];
  // prefix of sub-cluster counts of the chosen clusters (nc <= n_list <= 1024)
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (uint32_t j = 0; j < nc; ++j) {
      pre[j] = acc;
      uint32_t c = chosen_c[qi * n_list + j];
      acc += sub_off[c + 1] - sub_off[c];
    }
    pre[nc] = acc;
    base_sh = 0;
  }
  __syncthreads();
  const uint32_t row_len = pre[nc];
  const uint32_t* row = counts + qi * CS;
  for (uint32_t b = 0; b < row_len; b += kWG) {
    const uint32_t t = b + threadIdx.x;
    uint32_t sid = kNone, flag = 0;
    if (t < row_len) {
      uint32_t lo = 0, hi = nc;  // last j with pre[j] <= t
      while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= t) lo = mid; else hi = mid;
      }
      uint32_t c = chosen_c[qi * n_list + lo];
      sid = sub_off[c] + (t - pre[lo]);
      flag = row[sid] != 0;
    }
    uint32_t total;
    uint32_t pos = dev::wg_exclusive_scan(flag, tmp, total);
    if (flag) list_s[qi * cap + base_sh + pos] = sid;
    __syncthreads();
    if (threadIdx.x == 0) base_sh += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    nnz_s[qi] = base_sh;
    const uint64_t g = (q0 + qi) / G;
    atomicMax(&gstat[g * 8 + 1], (unsigned long long)base_sh);
    atomicMax(&gstat[g * 8 + 2], (unsigned long long)row_len);
  }
}

// A6 part 2: exact distances to the non-zero sub-centers, top ks = min(sub_lists, group max), with
// the group's table width supplying the 1e6 pad columns (Super-Grid-IVF_PQ.cu:885-896).  Output is
// sorted ascending (the reference's order here is RAFT-defined).  With a per-query budget array
// (PRECISION 1 with a budget model) the query keeps its own min(budget, sub_lists) nearest and no pads.
__global__ __launch_bounds__(256) void sub_select_kernel(
    const float* __restrict__ queries, uint64_t d, const float* __restrict__ sub_centers,
    uint32_t sub_lists, uint64_t G, uint64_t q0, uint32_t cap, const uint32_t* __restrict__ list_s,
    const uint32_t* __restrict__ nnz_s, const unsigned long long* __restrict__ gstat,
    const uint32_t* __restrict__ sub_budget, uint64_t* __restrict__ keys, uint32_t* __restrict__ chosen_s,
    uint32_t* __restrict__ n_chosen_s) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint64_t* sorted = reinterpret_cast<uint64_t*>(smem);
  float* stage = reinterpret_cast<float*>(smem);
  __shared__ float qv[512];
  __shared__ uint32_t rows[256];
  __shared__ uint32_t hist[2048];
  __shared__ uint64_t sh[8];
  const uint64_t qi = blockIdx.x;
  const float* q = queries + (q0 + qi) * d;
  for (uint32_t t = threadIdx.x; t < d; t += kWG) qv[t] = q[t];
  const uint32_t n = nnz_s[qi];
  uint64_t* gk = keys + qi * cap;
  __syncthreads();
  for (uint32_t b = 0; b < n; b += kWG) {
    const uint32_t j = b + threadIdx.x;
    rows[threadIdx.x] = j < n ? list_s[qi * cap + j] : kNone;
    __syncthreads();
    float dist = exact_tile(sub_centers, d, rows, qv, stage);
    if (j < n) gk[j] = dev::make_key(dist, j);
    __syncthreads();
  }
  __syncthreads();
  const uint64_t g = (q0 + qi) / G;
  uint32_t ks = (uint32_t)min<uint64_t>(sub_lists, gstat[g * 8 + 1]);
  uint64_t table_len = gstat[g * 8 + 2];
  if (sub_budget) {
    const uint32_t b = sub_budget[qi];
    ks = b == 0 || b > sub_lists ? sub_lists : b;
    table_len = n;
  }
  const uint32_t m = select_sorted(gk, n, ks, sorted, hist, sh);
  __syncthreads();
  const uint32_t keep = kept_with_pads(sorted, m, ks, table_len - n);
  for (uint32_t j = threadIdx.x; j < keep; j += kWG)
    chosen_s[qi * sub_lists + j] = list_s[qi * cap + dev::key_tie(sorted[j])];
  if (threadIdx.x == 0) n_chosen_s[qi] = keep;
}

// ------------------------------------------------------------------------------------------------
// A8/A9: per chosen sub-cluster, the Morton range [lower_bound(zlo), lower_bound(zhi+1)) inside its
// z-pool segment (zorder.cu:656-685).  One workgroup per query, one thread per chosen sub-cluster.
// Also sums C_q (candidates) and the in-box count of the chosen sub-clusters (A10 inputs).
__global__ __launch_bounds__(256) void ranges_kernel(
    const uint64_t* __restrict__ zcode, const uint32_t* __restrict__ seg_off,
    const uint32_t* __restrict__ counts, uint64_t CS, uint32_t sub_lists, uint64_t G, uint64_t q0,
    const uint32_t* __restrict__ chosen_s, const uint32_t* __restrict__ n_chosen_s,
    const uint64_t* __restrict__ zlo, const uint64_t* __restrict__ zhi, uint32_t* __restrict__ rstart,
    uint32_t* __restrict__ rpre, uint64_t* __restrict__ cand, uint64_t* __restrict__ real,
    unsigned long long* __restrict__ gstat) {
  __shared__ uint32_t tmp[8];
  __shared__ uint32_t base_sh;
  __shared__ unsigned long long real_sh;
  const uint64_t qi = blockIdx.x;
  const uint32_t n = n_chosen_s[qi];
  const uint64_t a = zlo[qi], b = zhi[qi] + 1;
  if (threadIdx.x == 0) {
    base_sh = 0;
    real_sh = 0;
  }
  __syncthreads();
  for (uint32_t j0 = 0; j0 < sub_lists; j0 += kWG) {
    const uint32_t j = j0 + threadIdx.x;
    uint32_t cnt = 0, start = 0;
    if (j < n) {
      const uint32_t s = chosen_s[qi * sub_lists + j];
      const uint32_t p0 = seg_off[s], p1 = seg_off[s + 1];
      uint32_t lo = p0, hi = p1;
      while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (zcode[mid] < a) lo = mid + 1; else hi = mid;
      }
      start = lo;
      uint32_t lo2 = lo, hi2 = p1;
      while (lo2 < hi2) {
        uint32_t mid = (lo2 + hi2) >> 1;
        if (zcode[mid] < b) lo2 = mid + 1; else hi2 = mid;
      }
      cnt = lo2 - lo;
      atomicAdd(&real_sh, (unsigned long long)counts[qi * CS + s]);
    }
    uint32_t total;
    uint32_t pos = dev::wg_exclusive_scan(cnt, tmp, total);
    if (j < sub_lists) {
      rstart[qi * sub_lists + j] = start;
      rpre[qi * (sub_lists + 1) + j] = base_sh + pos;
    }
    __syncthreads();
    if (threadIdx.x == 0) base_sh += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    rpre[qi * (sub_lists + 1) + sub_lists] = base_sh;
    cand[qi] = base_sh;
    real[qi] = real_sh;
    const uint64_t g = (q0 + qi) / G;
    atomicAdd(&gstat[g * 8 + 3], (unsigned long long)base_sh);
    atomicAdd(&gstat[g * 8 + 4], real_sh);
    atomicMax(&gstat[g * 8 + 5], (unsigned long long)base_sh);
  }
}

// f32 -> u64 as GCC emits it on x86-64 (the reference computes n_candies in host code).
__device__ __forceinline__ uint64_t f2u_host_dev(float x) {
  auto cvt = [](float v) -> int64_t {
    if (v != v || v >= 9223372036854775808.0f || v < -9223372036854775808.0f) return INT64_MIN;
    return (int64_t)v;
  };
  if (x < 9223372036854775808.0f) return (uint64_t)cvt(x);
  return (uint64_t)cvt(x - 9223372036854775808.0f) ^ 0x8000000000000000ull;
}

// A10 + buffer plan, one thread per query.  mode 0: every candidate reaches the next stage (the PQ
// top-k' keeps them all); mode 1: ADC scan + top-k'.  lsize = entries of the refine list,
// asize = ADC keys, chunks = adc_scan blocks.
__global__ void plan_kernel(uint64_t qn, uint64_t q0, uint64_t G, int l, uint64_t exps0, uint64_t topk,
                            const uint64_t* __restrict__ cand, const unsigned long long* __restrict__ gstat,
                            uint32_t* __restrict__ mode, uint64_t* __restrict__ kprime,
                            uint64_t* __restrict__ lsize, uint64_t* __restrict__ asize,
                            uint64_t* __restrict__ chunks, uint64_t* __restrict__ pad32,
                            uint64_t* __restrict__ padmax) {
  uint64_t qi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (qi >= qn) return;
  const uint64_t g = (q0 + qi) / G;
  const uint64_t Cq = cand[qi];
  const uint64_t W = gstat[g * 8 + 5];
  uint64_t kp;
  uint64_t p32 = 0, pmax = 0;
  if (exps0 <= 1) {
    kp = topk;  // query_prefilter path: label filter over all candidates, then PQ top-k
    mode[qi] = 2;
    lsize[qi] = Cq;
    asize[qi] = Cq;
    chunks[qi] = cdiv(Cq, kAdcChunk);
    kprime[qi] = kp;
    pad32[qi] = 0;
    padmax[qi] = 0;
    return;
  }
  if (l == 1) {
    kp = exps0 * topk;
    if (W >= kp) {
      p32 = kp - (Cq < kp ? Cq : kp);
    } else {
      p32 = W - Cq;
      pmax = kp - W;
    }
  } else {
    const float beta = (float)gstat[g * 8 + 3] / (float)gstat[g * 8 + 4];
    uint64_t x = f2u_host_dev((float)(exps0 * topk) * beta);
    kp = x < W ? x : W;
  }
  const bool adc = Cq > kp;
  mode[qi] = adc ? 1u : 0u;
  kprime[qi] = kp;
  lsize[qi] = adc ? kp : Cq;
  asize[qi] = adc ? Cq : 0;
  chunks[qi] = adc ? cdiv(Cq, kAdcChunk) : 0;
  pad32[qi] = p32;
  padmax[qi] = pmax;  // l > 1: set after the post-filter (n_candies - survivors)
}

// Exclusive scan of n u64 values in place (single workgroup), total to out[n].
__global__ __launch_bounds__(256) void scan_kernel(uint64_t* __restrict__ v, uint64_t n) {
  __shared__ unsigned long long carry_sh;
  __shared__ unsigned long long part[256];
  if (threadIdx.x == 0) carry_sh = 0;
  __syncthreads();
  for (uint64_t b = 0; b < n; b += 256) {
    uint64_t i = b + threadIdx.x;
    unsigned long long x = i < n ? v[i] : 0;
    part[threadIdx.x] = x;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
      unsigned long long y = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0;
      __syncthreads();
      part[threadIdx.x] += y;
      __syncthreads();
    }
    unsigned long long incl = part[threadIdx.x];
    if (i < n) v[i] = carry_sh + incl - x;
    __syncthreads();
    if (threadIdx.x == 255) carry_sh += incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) v[n] = carry_sh;
}

// Position in the z-pool of candidate column `col` of query row (rstart/rpre of that query).
__device__ __forceinline__ uint32_t col_to_pos(const uint32_t* rs, const uint32_t* rp, uint32_t n,
                                               uint32_t col) {
  uint32_t lo = 0, hi = n;  // last j with rp[j] <= col
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (rp[mid] <= col) lo = mid; else hi = mid;
  }
  return rs[lo] + (col - rp[lo]);
}

__device__ __forceinline__ bool label_pass(const float* __restrict__ lab, int l, uint32_t pos,
                                           const float* rg) {
  for (int i = 0; i < l; ++i) {
    float v = lab[(uint64_t)pos * l + i];
    if (v < rg[2 * i] || v > rg[2 * i + 1]) return false;
  }
  return true;
}

// Mode 0 / 2: copy every candidate (l == 1) or every candidate passing the exact label predicate
// (l > 1 and the query_prefilter path) into the refine list.  One workgroup per query.
__global__ __launch_bounds__(256) void keep_all_kernel(
    uint64_t qn, int l, uint32_t sub_lists, const uint32_t* __restrict__ mode,
    const uint32_t* __restrict__ n_chosen_s, const uint32_t* __restrict__ rstart,
    const uint32_t* __restrict__ rpre, const float* __restrict__ lab, const float* __restrict__ ranges,
    const uint64_t* __restrict__ loff, uint32_t* __restrict__ L, uint32_t* __restrict__ nL) {
  __shared__ uint32_t tmp[8];
  const uint64_t qi = blockIdx.x;
  if (mode[qi] == 1) return;
  const bool filt = (l > 1) || mode[qi] == 2;
  const uint32_t nr = n_chosen_s[qi];
  const uint32_t* rs = rstart + qi * sub_lists;
  const uint32_t* rp = rpre + qi * (sub_lists + 1);
  const uint32_t Cq = rp[nr];
  uint32_t* out = L + loff[qi];
  if (!filt) {
    for (uint32_t j = 0; j < nr; ++j) {
      const uint32_t p0 = rs[j], n = rp[j + 1] - rp[j];
      for (uint32_t e = threadIdx.x; e < n; e += kWG) out[rp[j] + e] = p0 + e;
    }
    if (threadIdx.x == 0) nL[qi] = Cq;
    return;
  }
  // order-preserving compaction in candidate-column order (filter.cu:146-164)
  float rg[2 * kMaxDims];
  for (int i = 0; i < 2 * l; ++i) rg[i] = ranges[qi * 2 * l + i];
  uint32_t base = 0;
  for (uint32_t b = 0; b < Cq; b += kWG) {
    const uint32_t col = b + threadIdx.x;
    uint32_t pos = 0, flag = 0;
    if (col < Cq) {
      pos = col_to_pos(rs, rp, nr, col);
      flag = label_pass(lab, l, pos, rg) ? 1u : 0u;
    }
    uint32_t total;
    const uint32_t at = dev::wg_exclusive_scan(flag, tmp, total);
    if (flag) out[base + at] = pos;
    base += total;
  }
  if (threadIdx.x == 0) nL[qi] = base;
}

// A11: ADC distances (product-quantization.cu:37-77) written as unique keys (distance bits, column).
// Mode 1: the query's Morton-range candidates, column = candidate column.  Mode 2 (EXPS0 <= 1): the
// label-filtered refine list, column = list index.  One block per kAdcChunk items of a query; each
// thread owns kAdcChunk/256 items and keeps their partial sums in registers while the LUT
// (product-quantization.cu:4-35, an fma chain over pq_len) passes through LDS in phases of
// kLutPhaseBytes, so the sum over m stays sequential for any pq_dim.
constexpr uint32_t kLutPhaseBytes = 48 * 1024;
constexpr int kAdcPer = kAdcChunk / kWG;

// LUT entry types: f32 (reference-exact) or f16 (PRECISION 1); sums always accumulate in f32.
__device__ __forceinline__ float lutf(float v) { return v; }
__device__ __forceinline__ float lutf(__half v) { return __half2float(v); }
__device__ __forceinline__ void lut_store(float* p, float v) { *p = v; }
__device__ __forceinline__ void lut_store(__half* p, float v) { *p = __float2half(v); }

template <typename LT>
__global__ __launch_bounds__(256) void adc_scan_kernel(
    uint64_t qn, uint64_t q0, const float* __restrict__ queries, uint64_t d, uint64_t M, uint64_t K,
    uint64_t pq_len, const float* __restrict__ pq_centers, const uint8_t* __restrict__ codes,
    uint32_t sub_lists, const uint32_t* __restrict__ mode, const uint32_t* __restrict__ n_chosen_s,
    const uint32_t* __restrict__ rstart, const uint32_t* __restrict__ rpre, const uint64_t* __restrict__ cand,
    const uint32_t* __restrict__ L, const uint64_t* __restrict__ loff, const uint32_t* __restrict__ nL,
    const uint64_t* __restrict__ coff, const uint64_t* __restrict__ aoff, uint64_t* __restrict__ akeys) {
  extern __shared__ __align__(16) unsigned char smem[];
  LT* lut = reinterpret_cast<LT*>(smem);  // [mp][K]
  // which query: last qi with coff[qi] <= blockIdx.x
  const uint64_t blk = blockIdx.x;
  uint64_t lo = 0, hi = qn;
  while (hi - lo > 1) {
    uint64_t mid = (lo + hi) >> 1;
    if (coff[mid] <= blk) lo = mid; else hi = mid;
  }
  const uint64_t qi = lo;
  const bool from_list = mode[qi] == 2;
  const uint64_t n_items = from_list ? nL[qi] : cand[qi];
  const uint64_t c0 = (blk - coff[qi]) * kAdcChunk;
  if (c0 >= n_items) return;
  const uint64_t c1 = c0 + kAdcChunk < n_items ? c0 + kAdcChunk : n_items;
  const float* q = queries + (q0 + qi) * d;
  const uint32_t nr = n_chosen_s[qi];
  const uint32_t* rs = rstart + qi * sub_lists;
  const uint32_t* rp = rpre + qi * (sub_lists + 1);
  const uint32_t* list = L + loff[qi];
  uint32_t pos[kAdcPer];
  float part[kAdcPer];
#pragma unroll
  for (int i = 0; i < kAdcPer; ++i) {
    const uint64_t col = c0 + threadIdx.x + (uint64_t)i * kWG;
    part[i] = 0.0f;
    pos[i] = 0;
    if (col < c1) pos[i] = from_list ? list[col] : col_to_pos(rs, rp, nr, (uint32_t)col);
  }
  uint64_t mp = kLutPhaseBytes / (K * sizeof(LT));
  if (mp > 4) mp &= ~3ull;
  const bool words = (M % 4) == 0 && (mp % 4) == 0;
  for (uint64_t m0 = 0; m0 < M; m0 += mp) {
    const uint64_t m1 = m0 + mp < M ? m0 + mp : M;
    for (uint64_t e = threadIdx.x; e < (m1 - m0) * K; e += kWG) {
      const uint64_t m = m0 + e / K, c = e % K;
      float ans = 0.0f;
      for (uint64_t t = 0; t < pq_len; ++t) {
        if (m * pq_len + t >= d) break;
        float diff = pq_centers[m * pq_len * K + c * pq_len + t] - q[m * pq_len + t];
        ans = fmaf(diff, diff, ans);
      }
      lut_store(&lut[e], ans);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kAdcPer; ++i) {
      const uint64_t col = c0 + threadIdx.x + (uint64_t)i * kWG;
      if (col >= c1) continue;
      const uint8_t* cd = codes + (uint64_t)pos[i] * M;
      float sum = part[i];
      if (words) {
        for (uint64_t m = m0; m < m1; m += 4) {
          const uint32_t wv = *reinterpret_cast<const uint32_t*>(cd + m);
          const LT* lr = lut + (m - m0) * K;
          sum += lutf(lr[wv & 0xFF]);
          sum += lutf(lr[K + ((wv >> 8) & 0xFF)]);
          sum += lutf(lr[2 * K + ((wv >> 16) & 0xFF)]);
          sum += lutf(lr[3 * K + (wv >> 24)]);
        }
      } else {
        for (uint64_t m = m0; m < m1; ++m) sum += lutf(lut[(m - m0) * K + cd[m]]);
      }
      part[i] = sum;
    }
    __syncthreads();
  }
  uint64_t* out = akeys + aoff[qi];
#pragma unroll
  for (int i = 0; i < kAdcPer; ++i) {
    const uint64_t col = c0 + threadIdx.x + (uint64_t)i * kWG;
    if (col < c1) out[col] = dev::make_key(part[i], (uint32_t)col);
  }
}

// A11 selection + A12 post-filter for mode-1 queries: the k' smallest (pq distance, column) keys
// (select_k, product-quantization.cu:264), then for l > 1 the exact label filter.
__global__ __launch_bounds__(256) void adc_select_kernel(
    uint64_t qn, int l, uint32_t sub_lists, const uint32_t* __restrict__ mode,
    const uint64_t* __restrict__ kprime, const uint64_t* __restrict__ cand,
    const uint32_t* __restrict__ n_chosen_s, const uint32_t* __restrict__ rstart,
    const uint32_t* __restrict__ rpre, const float* __restrict__ lab, const float* __restrict__ ranges,
    const uint64_t* __restrict__ aoff, uint64_t* __restrict__ akeys_sel_scratch,
    const uint64_t* __restrict__ akeys, const uint64_t* __restrict__ loff, uint32_t* __restrict__ L,
    uint32_t* __restrict__ nL) {
  __shared__ uint32_t hist[2048];
  __shared__ uint64_t sh[8];
  __shared__ uint32_t cnt_sh;
  const uint64_t qi = blockIdx.x;
  if (mode[qi] != 1) return;
  const uint32_t Cq = (uint32_t)cand[qi];
  const uint32_t kp = (uint32_t)kprime[qi];
  if (kp == 0) {
    if (threadIdx.x == 0) nL[qi] = 0;
    return;
  }
  uint64_t* sel = akeys_sel_scratch + loff[qi];  // kp keys
  dev::wg_radix_select(akeys + aoff[qi], Cq, kp, sel, hist, sh);
  const uint32_t nr = n_chosen_s[qi];
  const uint32_t* rs = rstart + qi * sub_lists;
  const uint32_t* rp = rpre + qi * (sub_lists + 1);
  float rg[2 * kMaxDims];
  for (int i = 0; i < 2 * l; ++i) rg[i] = ranges[qi * 2 * l + i];
  if (threadIdx.x == 0) cnt_sh = 0;
  __syncthreads();
  uint32_t* out = L + loff[qi];
  for (uint32_t e = threadIdx.x; e < kp; e += kWG) {
    const uint32_t pos = col_to_pos(rs, rp, nr, dev::key_tie(sel[e]));
    if (l == 1) {
      out[e] = pos;
    } else if (label_pass(lab, l, pos, rg)) {
      out[atomicAdd(&cnt_sh, 1u)] = pos;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) nL[qi] = l == 1 ? kp : cnt_sh;
}

// A13: exact FP32 distances of the refine list, top-k by (distance, id) with the reference's pad
// columns (refine.cu:78-119), or for the query_prefilter path the top-k by (pq distance, column).
__global__ __launch_bounds__(256) void final_kernel(
    uint64_t qn, uint64_t q0, const float* __restrict__ queries, uint64_t d, const float* __restrict__ vec,
    const uint32_t* __restrict__ zid, int l, uint64_t topk, const uint32_t* __restrict__ mode,
    const uint64_t* __restrict__ kprime, const uint64_t* __restrict__ loff, const uint32_t* __restrict__ L,
    const uint32_t* __restrict__ nL, const uint64_t* __restrict__ pad32, const uint64_t* __restrict__ padmax,
    uint64_t* __restrict__ keys, uint64_t* __restrict__ akeys, const uint64_t* __restrict__ aoff,
    float* __restrict__ out_dist, uint64_t* __restrict__ out_ids) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint64_t* sorted = reinterpret_cast<uint64_t*>(smem);
  float* stage = reinterpret_cast<float*>(smem);
  __shared__ float qv[512];
  __shared__ uint32_t rows[256];
  __shared__ uint32_t hist[2048];
  __shared__ uint64_t sh[8];
  const uint64_t qi = blockIdx.x;
  const float* q = queries + (q0 + qi) * d;
  for (uint32_t t = threadIdx.x; t < d; t += kWG) qv[t] = q[t];
  const uint32_t n = nL[qi];
  const uint32_t* list = L + loff[qi];
  uint64_t* gk = keys + loff[qi];
  float* od = out_dist + (q0 + qi) * topk;
  uint64_t* oi = out_ids + (q0 + qi) * topk;
  __syncthreads();
  if (mode[qi] == 2) {
    // query_prefilter path: adc_scan_kernel left the PQ keys of the filtered list in akeys
    gk = akeys + aoff[qi];
  } else {
    for (uint32_t b = 0; b < n; b += kWG) {
      const uint32_t j = b + threadIdx.x;
      rows[threadIdx.x] = j < n ? list[j] : kNone;
      __syncthreads();
      float dist = exact_tile(vec, d, rows, qv, stage);
      if (j < n) gk[j] = dev::make_key(dist, zid[list[j]]);
      __syncthreads();
    }
  }
  __syncthreads();
  const uint32_t k = (uint32_t)topk;
  const uint32_t m = select_sorted(gk, n, k, sorted, hist, sh);
  __syncthreads();
  if (mode[qi] == 2) {
    for (uint32_t j = threadIdx.x; j < k; j += kWG) {
      if (j < m) {
        od[j] = dev::key_value(sorted[j]);
        oi[j] = zid[list[dev::key_tie(sorted[j])]];
      } else {
        od[j] = FLT_MAX;
        oi[j] = kU64Max;
      }
    }
    return;
  }
  // pads of the refine list: (1e6, 2^32) x p32 then (1e6, UINT64_MAX) x pmax
  uint64_t p32 = pad32[qi], pmax = padmax[qi];
  if (l > 1) pmax = kprime[qi] > n ? kprime[qi] - n : 0;
  if (threadIdx.x == 0) {
    uint32_t n_le = 0;
    while (n_le < m && dev::key_value(sorted[n_le]) <= kExactPad) ++n_le;
    uint32_t j = 0, r = 0;
    for (; j < k && r < n_le; ++j, ++r) {
      od[j] = dev::key_value(sorted[r]);
      oi[j] = dev::key_tie(sorted[r]);
    }
    for (; j < k && p32 > 0; ++j, --p32) {
      od[j] = kExactPad;
      oi[j] = kIdPad;
    }
    for (; j < k && pmax > 0; ++j, --pmax) {
      od[j] = kExactPad;
      oi[j] = kU64Max;
    }
    for (; j < k && r < m; ++j, ++r) {
      od[j] = dev::key_value(sorted[r]);
      oi[j] = dev::key_tie(sorted[r]);
    }
    for (; j < k; ++j) {
      od[j] = kExactPad;
      oi[j] = kU64Max;
    }
  }
}

}  // namespace

// ================================================================================================
// Host orchestration
struct Index::Work {
  DevBuf ranges, boxes, box_lo, box_w, zlo, zhi, zlo_cnt, err;
  DevBuf counts, list_c, nnz_c, chosen_c, n_chosen_c, list_s, nnz_s, chosen_s, n_chosen_s;
  DevBuf keys, gstat, rstart, rpre, cand, real;
  DevBuf mode, kprime, lsize, asize, chunks, pad32, padmax, nL;
  DevBuf L, Lkeys, akeys;
  DevBuf S, feat, budget_T, budget_S;  // PRECISION 1: coarse scores, model features, per-query budgets
  uint64_t* pinned = nullptr;
  hipEvent_t ev[16];
  bool ev_init = false;
};

int find_max_factor(uint64_t a, uint64_t b) {
  if (b <= a) return (int)b;
  int best = 1;
  for (int i = 1; (double)i <= std::sqrt((double)b); ++i) {
    if (b % (uint64_t)i) continue;
    if ((uint64_t)i < a) best = std::max(best, i);
    int pair = (int)(b / (uint64_t)i);
    if ((uint64_t)pair < a) best = std::max(best, pair);
  }
  return best;
}

Index::Index(const Config& cfg, const FilterConfig& fc, int device) : cfg_(cfg), fc_(fc), device_(device) {
  w_ = new Work();
  HDG_HIP(hipSetDevice(device_));
  HDG_HIP(hipHostMalloc(reinterpret_cast<void**>(&w_->pinned), 64 * sizeof(uint64_t), hipHostMallocDefault));
  std::vector<int> ft(fc_.filter_type.begin(), fc_.filter_type.end());
  std::vector<float> sv(fc_.shift_val.begin(), fc_.shift_val.end());
  std::vector<float> mp = fc_.flat_map();
  if (mp.size() < 2) mp.resize(2, 0.0f);  // the kernel reads map[0..1] after flagging a bad label
  f_map_len_ = fc_.flat_map().size();
  std::vector<int> dv(fc_.div_value.begin(), fc_.div_value.end());
  upload(f_type_, ft, nullptr);
  upload(f_shift_, sv, nullptr);
  upload(f_map_, mp, nullptr);
  upload(f_div_, dv, nullptr);
}

Index::~Index() {
  budget_free();
  if (w_) {
    (void)hipSetDevice(device_);
    if (w_->pinned) (void)hipHostFree(w_->pinned);
    if (w_->ev_init)
      for (auto& e : w_->ev) (void)hipEventDestroy(e);
    delete w_;
  }
}

void Index::debug_coarse_scores(const float* queries, uint64_t nq, float* out_host, hipStream_t s) {
  HDG_HIP(hipSetDevice(device_));
  if (!st_.built) fail(kBadState, "query before build");
  DevBuf S;
  S.reserve(nq * st_.C * 4);
  launch_coarse_gemm(queries, 0, nq, st_.d, st_.Dp, st_.centers_bf16.as<uint16_t>(), st_.center_norm.as<float>(),
                     (uint32_t)st_.C, S.as<float>(), s);
  HDG_HIP(hipGetLastError());
  HDG_HIP(hipMemcpyAsync(out_host, S.data(), nq * st_.C * 4, hipMemcpyDeviceToHost, s));
  HDG_HIP(hipStreamSynchronize(s));
}

void Index::query(const float* queries, uint64_t nq, uint64_t d, const float* qlabels, uint64_t qcols,
                  float* out_dist, uint64_t* out_ids, hipStream_t s) {
  HDG_HIP(hipSetDevice(device_));
  if (!st_.built) fail(kBadState, "query before build");
  if (d != st_.d) fail(kInvalidArgument, "query dimension differs from the index");
  if (nq == 0) return;
  if (!queries || !qlabels || !out_dist || !out_ids) fail(kInvalidArgument, "null query/output pointer");
  const int l = st_.l;
  if (qcols < (uint64_t)l) fail(kInvalidArgument, "query label columns < l");
  for (int i = 0; i < l; ++i)
    if (fc_.filter_type[i] == 3 && qcols < 2 * (uint64_t)l)
      fail(kInvalidArgument, "explicit-range filter needs 2*l query label columns");
  uint64_t G = cfg_.parity_batch ? cfg_.parity_batch : (uint64_t)find_max_factor(125, nq);
  uint64_t chunk = std::max<uint64_t>(G, (cfg_.query_chunk / G) * G);
  stats_ = QueryStats();
  stats_.nq = nq;
  stats_.cand.resize(nq);
  stats_.real.resize(nq);
  stats_.refined.resize(nq);
  stats_.scored.resize(nq);
  stats_.adc.resize(nq);
  stats_.kprime.resize(nq);
  stage_ms.clear();
  for (uint64_t q0 = 0; q0 < nq; q0 += chunk) {
    uint64_t qn = std::min(chunk, nq - q0);
    query_chunk(queries, q0, qn, qlabels, qcols, G, out_dist, out_ids, s);
  }
}

void Index::query_chunk(const float* queries, uint64_t q0, uint64_t qn, const float* qlabels, uint64_t qcols,
                        uint64_t G, float* out_dist, uint64_t* out_ids, hipStream_t s) {
  Work& w = *w_;
  const IndexState& S = st_;
  const GridShape& g = S.grid;
  const int l = S.l;
  const uint64_t d = S.d, C = S.C, CS = S.CS;
  const uint32_t n_list = (uint32_t)std::min<uint64_t>(cfg_.n_list, C);
  const uint32_t sub_lists = (uint32_t)cfg_.sub_lists;
  const uint64_t topk = cfg_.topk;
  if (d > 512) fail(kUnsupported, "vector dimension above 512");
  if (C > (uint64_t)dev::kSortCap) fail(kUnsupported, "clusters above 8192");
  if (cfg_.n_list > 1024) fail(kUnsupported, "n_list above 1024");
  if (topk > (uint64_t)dev::kSortCap || topk == 0) fail(kUnsupported, "topk must be in [1, 8192]");
  if (sub_lists == 0 || sub_lists > (uint32_t)dev::kSortCap) fail(kUnsupported, "sub_lists must be in [1, 8192]");
  const uint32_t cap = (uint32_t)std::min<uint64_t>(CS, (uint64_t)n_list * S.max_subs_per_cluster);
  const uint64_t ng = cdiv(q0 + qn, G) - q0 / G;  // groups touched by this chunk
  const uint64_t g_base = q0 / G;
  const Dims D = make_dims(g);

  // workspace
  w.ranges.reserve(qn * 2 * l * 4);
  w.boxes.reserve(qn * 2 * l * 4);
  w.box_lo.reserve((uint64_t)std::max(1, g.n_valid) * qn * 4);
  w.box_w.reserve((uint64_t)std::max(1, g.n_valid) * qn * 4);
  w.zlo.reserve(qn * 8);
  w.zhi.reserve(qn * 8);
  w.zlo_cnt.reserve(qn * 8);
  w.err.reserve(4);
  w.counts.reserve(qn * CS * 4);
  w.list_c.reserve(qn * C * 4);
  w.nnz_c.reserve(qn * 4);
  w.chosen_c.reserve(qn * n_list * 4);
  w.n_chosen_c.reserve(qn * 4);
  w.list_s.reserve(qn * cap * 4);
  w.nnz_s.reserve(qn * 4);
  w.chosen_s.reserve(qn * sub_lists * 4);
  w.n_chosen_s.reserve(qn * 4);
  w.keys.reserve(qn * std::max<uint64_t>(C, cap) * 8);
  w.gstat.reserve((ng + g_base + 1) * 8 * 8);
  w.rstart.reserve(qn * sub_lists * 4);
  w.rpre.reserve(qn * (sub_lists + 1) * 4);
  w.cand.reserve(qn * 8);
  w.real.reserve(qn * 8);
  w.mode.reserve(qn * 4);
  w.kprime.reserve(qn * 8);
  w.lsize.reserve((qn + 1) * 8);
  w.asize.reserve((qn + 1) * 8);
  w.chunks.reserve((qn + 1) * 8);
  w.pad32.reserve(qn * 8);
  w.padmax.reserve(qn * 8);
  w.nL.reserve(qn * 4);
  unsigned long long* gstat = w.gstat.as<unsigned long long>() + g_base * 8;
  HDG_HIP(hipMemsetAsync(gstat, 0, ng * 8 * 8, s));
  HDG_HIP(hipMemsetAsync(w.err.data(), 0, 4, s));
  HDG_HIP(hipMemsetAsync(w.chosen_s.data(), 0xFF, qn * sub_lists * 4, s));

  // dynamic LDS of the per-query kernels: the sort buffer, aliased by the 256 x 33 distance stage
  const size_t sort_lds = std::max<size_t>(dev::kSortCap * sizeof(uint64_t), 256 * 33 * sizeof(float));
  if (!w.ev_init) {
    for (auto& e : w.ev) HDG_HIP(hipEventCreate(&e));
    w.ev_init = true;
  }
  int evn = 0;
  auto mark = [&]() { HDG_HIP(hipEventRecord(w.ev[evn++], s)); };
  mark();
  hipLaunchKernelGGL(prep_kernel, dim3((unsigned)cdiv(qn, 128)), dim3(128), 0, s, qlabels, qcols, q0, qn, G, D,
                     f_type_.as<int>(), f_shift_.as<float>(), f_map_.as<float>(), f_map_len_, f_div_.as<int>(),
                     w.ranges.as<float>(), w.boxes.as<int>(), w.box_lo.as<uint32_t>(), w.box_w.as<uint32_t>(),
                     w.zlo.as<uint64_t>(), w.zhi.as<uint64_t>(), w.zlo_cnt.as<uint64_t>(), w.err.as<uint32_t>());
  HDG_HIP(hipGetLastError());
  mark();
  const bool mixed = cfg_.precision == 1;
  // PRECISION 1 with PROBE_CLUSTERS > 0: BF16 coarse top-T, counts of those clusters only (lazy);
  // PRECISION 1 with PROBE_CLUSTERS = 0: exact full counts, BF16 MFMA scores for cluster and sub-cluster choice.
  const bool lazy = mixed && cfg_.probe_clusters > 0;
  uint32_t* budget_T = nullptr;
  uint32_t* budget_S = nullptr;
  if (mixed && cfg_.budget_model == 1) {
    if (!budget_model_ready()) fail(kBadState, "BUDGET_MODEL = 1 but no budget model is loaded");
    w.budget_T.reserve(qn * 4);
    w.budget_S.reserve(qn * 4);
    budget_T = w.budget_T.as<uint32_t>();
    budget_S = w.budget_S.as<uint32_t>();
  }
  if (mixed) {
    w.S.reserve(qn * std::max<uint64_t>(C, CS) * 4);
    w.feat.reserve(qn * kFeatDists * 4);
  }
  auto launch_full_count = [&]() {
    dim3 grid((unsigned)cdiv(qn, 256), (unsigned)cdiv(CS, kCountSegChunk));
    const uint32_t* cells = S.cells.as<uint32_t>();
    const uint32_t* so = S.seg_off.as<uint32_t>();
    switch (g.n_valid) {
#define HDG_COUNT_CASE(NV)                                                                              \
  case NV:                                                                                              \
    hipLaunchKernelGGL(count_kernel<NV>, grid, dim3(256), 0, s, cells, S.N, S.zcode.as<uint64_t>(), so, CS,     \
                       w.box_lo.as<uint32_t>(), w.box_w.as<uint32_t>(), w.zlo_cnt.as<uint64_t>(),             \
                       w.zhi.as<uint64_t>(), qn, w.counts.as<uint32_t>());                                    \
    break;
      HDG_COUNT_CASE(1)
      HDG_COUNT_CASE(2)
      HDG_COUNT_CASE(3)
      HDG_COUNT_CASE(4)
      HDG_COUNT_CASE(5)
      HDG_COUNT_CASE(6)
      HDG_COUNT_CASE(7)
      HDG_COUNT_CASE(8)
#undef HDG_COUNT_CASE
      default:
        fail(kBadState, "no valid label dims");
    }
    HDG_HIP(hipGetLastError());
  };
  auto budgets = [&]() {
    capture_features(q0, qn, l, w.ranges.as<float>(), w.box_w.as<uint32_t>(), g.n_valid, w.feat.as<float>(), s);
    if (budget_T)
      run_budget_model(qn, l, w.ranges.as<float>(), w.box_lo.as<uint32_t>(), w.box_w.as<uint32_t>(), g.n_valid,
                       w.feat.as<float>(), w.n_chosen_c.as<uint32_t>(), budget_T, budget_S, s);
  };
  if (lazy) {
    launch_coarse_gemm(queries, q0, qn, d, S.Dp, S.centers_bf16.as<uint16_t>(), S.center_norm.as<float>(),
                       (uint32_t)C, w.S.as<float>(), s);
    HDG_HIP(hipGetLastError());
    mark();
    launch_coarse_topk(queries, q0, qn, d, w.S.as<float>(), (uint32_t)C, n_list, (uint32_t)cfg_.probe_clusters,
                       nullptr, w.chosen_c.as<uint32_t>(), w.n_chosen_c.as<uint32_t>(), w.feat.as<float>(), s);
    HDG_HIP(hipGetLastError());
    budgets();
    mark();
    launch_budget_count(g.n_valid, S.cells.as<uint32_t>(), S.N, S.zcode.as<uint64_t>(), S.seg_off.as<uint32_t>(),
                        S.sub_off.as<uint32_t>(), CS, w.box_lo.as<uint32_t>(), w.box_w.as<uint32_t>(),
                        w.zlo_cnt.as<uint64_t>(), w.zhi.as<uint64_t>(), qn, n_list, w.chosen_c.as<uint32_t>(),
                        w.n_chosen_c.as<uint32_t>(), w.counts.as<uint32_t>(), s);
    HDG_HIP(hipGetLastError());
    mark();
  } else {
    launch_full_count();
    mark();
    hipLaunchKernelGGL(cluster_nnz_kernel, dim3((unsigned)qn), dim3(256), 0, s, w.counts.as<uint32_t>(), CS,
                       S.sub_off.as<uint32_t>(), (uint32_t)C, G, q0, w.list_c.as<uint32_t>(), w.nnz_c.as<uint32_t>(),
                       w.gstat.as<unsigned long long>());
    HDG_HIP(hipGetLastError());
    mark();
    if (!mixed) {
      hipLaunchKernelGGL(coarse_select_kernel, dim3((unsigned)qn), dim3(256), sort_lds, s, queries, d,
                         S.centers.as<float>(), (uint32_t)C, n_list, G, q0, w.list_c.as<uint32_t>(),
                         w.nnz_c.as<uint32_t>(), w.gstat.as<unsigned long long>(), w.keys.as<uint64_t>(),
                         w.chosen_c.as<uint32_t>(), w.n_chosen_c.as<uint32_t>());
    } else {
      launch_coarse_gemm(queries, q0, qn, d, S.Dp, S.centers_bf16.as<uint16_t>(), S.center_norm.as<float>(),
                         (uint32_t)C, w.S.as<float>(), s);
      launch_scored_select(queries, q0, qn, d, w.list_c.as<uint32_t>(), (uint32_t)C, w.nnz_c.as<uint32_t>(),
                           w.S.as<float>(), C, n_list, nullptr, n_list, w.chosen_c.as<uint32_t>(),
                           w.n_chosen_c.as<uint32_t>(), w.feat.as<float>(), w.keys.as<uint64_t>(), s);
      budgets();
    }
    HDG_HIP(hipGetLastError());
    mark();
  }
  hipLaunchKernelGGL(sub_nnz_kernel, dim3((unsigned)qn), dim3(256), 0, s, w.counts.as<uint32_t>(), CS,
                     S.sub_off.as<uint32_t>(), n_list, G, q0, w.chosen_c.as<uint32_t>(), w.n_chosen_c.as<uint32_t>(),
                     cap, w.list_s.as<uint32_t>(), w.nnz_s.as<uint32_t>(), w.gstat.as<unsigned long long>());
  HDG_HIP(hipGetLastError());
  mark();
  if (mixed && !lazy) {
    launch_coarse_gemm(queries, q0, qn, d, S.Dp, S.sub_centers_bf16.as<uint16_t>(), S.sub_center_norm.as<float>(),
                       (uint32_t)CS, w.S.as<float>(), s);
    launch_scored_select(queries, q0, qn, d, w.list_s.as<uint32_t>(), cap, w.nnz_s.as<uint32_t>(), w.S.as<float>(),
                         CS, sub_lists, budget_S, sub_lists, w.chosen_s.as<uint32_t>(), w.n_chosen_s.as<uint32_t>(),
                         nullptr, w.keys.as<uint64_t>(), s);
  } else {
    hipLaunchKernelGGL(sub_select_kernel, dim3((unsigned)qn), dim3(256), sort_lds, s, queries, d,
                       S.sub_centers.as<float>(), sub_lists, G, q0, cap, w.list_s.as<uint32_t>(),
                       w.nnz_s.as<uint32_t>(), w.gstat.as<unsigned long long>(), budget_S, w.keys.as<uint64_t>(),
                       w.chosen_s.as<uint32_t>(), w.n_chosen_s.as<uint32_t>());
  }
  HDG_HIP(hipGetLastError());
  mark();
  hipLaunchKernelGGL(ranges_kernel, dim3((unsigned)qn), dim3(256), 0, s, S.zcode.as<uint64_t>(),
                     S.seg_off.as<uint32_t>(), w.counts.as<uint32_t>(), CS, sub_lists, G, q0,
                     w.chosen_s.as<uint32_t>(), w.n_chosen_s.as<uint32_t>(), w.zlo.as<uint64_t>(),
                     w.zhi.as<uint64_t>(), w.rstart.as<uint32_t>(), w.rpre.as<uint32_t>(), w.cand.as<uint64_t>(),
                     w.real.as<uint64_t>(), w.gstat.as<unsigned long long>());
  HDG_HIP(hipGetLastError());
  mark();
  hipLaunchKernelGGL(plan_kernel, dim3((unsigned)cdiv(qn, 128)), dim3(128), 0, s, qn, q0, G, l, cfg_.exps0,
                     topk, w.cand.as<uint64_t>(), w.gstat.as<unsigned long long>(), w.mode.as<uint32_t>(),
                     w.kprime.as<uint64_t>(), w.lsize.as<uint64_t>(), w.asize.as<uint64_t>(),
                     w.chunks.as<uint64_t>(), w.pad32.as<uint64_t>(), w.padmax.as<uint64_t>());
  HDG_HIP(hipGetLastError());
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(256), 0, s, w.lsize.as<uint64_t>(), qn);
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(256), 0, s, w.asize.as<uint64_t>(), qn);
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(256), 0, s, w.chunks.as<uint64_t>(), qn);
  HDG_HIP(hipGetLastError());
  // the one host synchronisation of a chunk: sizes of the data-dependent buffers
  HDG_HIP(hipMemcpyAsync(&w.pinned[0], w.lsize.as<uint64_t>() + qn, 8, hipMemcpyDeviceToHost, s));
  HDG_HIP(hipMemcpyAsync(&w.pinned[1], w.asize.as<uint64_t>() + qn, 8, hipMemcpyDeviceToHost, s));
  HDG_HIP(hipMemcpyAsync(&w.pinned[2], w.chunks.as<uint64_t>() + qn, 8, hipMemcpyDeviceToHost, s));
  HDG_HIP(hipMemcpyAsync(&w.pinned[3], w.err.data(), 4, hipMemcpyDeviceToHost, s));
  HDG_HIP(hipStreamSynchronize(s));
  mark();
  if ((uint32_t)w.pinned[3] & 1u) fail(kInvalidArgument, "map filter label outside interval_map");
  const uint64_t totL = w.pinned[0], totA = w.pinned[1], totChunks = w.pinned[2];
  w.L.reserve(std::max<uint64_t>(totL, 1) * 4);
  w.Lkeys.reserve(std::max<uint64_t>(totL, 1) * 8);
  w.akeys.reserve(std::max<uint64_t>(totA, 1) * 8);
  hipLaunchKernelGGL(keep_all_kernel, dim3((unsigned)qn), dim3(256), 0, s, qn, l, sub_lists, w.mode.as<uint32_t>(),
                     w.n_chosen_s.as<uint32_t>(), w.rstart.as<uint32_t>(), w.rpre.as<uint32_t>(), S.lab.as<float>(),
                     w.ranges.as<float>(), w.lsize.as<uint64_t>(), w.L.as<uint32_t>(), w.nL.as<uint32_t>());
  HDG_HIP(hipGetLastError());
  mark();
  if (totChunks > 0) {
    auto kern = mixed ? adc_scan_kernel<__half> : adc_scan_kernel<float>;
    hipLaunchKernelGGL(kern, dim3((unsigned)totChunks), dim3(256), kLutPhaseBytes, s, qn, q0, queries,
                       d, S.M, S.K, S.pq_len, S.pq_centers.as<float>(), S.codes.as<uint8_t>(), sub_lists,
                       w.mode.as<uint32_t>(), w.n_chosen_s.as<uint32_t>(), w.rstart.as<uint32_t>(),
                       w.rpre.as<uint32_t>(), w.cand.as<uint64_t>(), w.L.as<uint32_t>(), w.lsize.as<uint64_t>(),
                       w.nL.as<uint32_t>(), w.chunks.as<uint64_t>(), w.asize.as<uint64_t>(), w.akeys.as<uint64_t>());
    HDG_HIP(hipGetLastError());
    hipLaunchKernelGGL(adc_select_kernel, dim3((unsigned)qn), dim3(256), 0, s, qn, l, sub_lists,
                       w.mode.as<uint32_t>(), w.kprime.as<uint64_t>(), w.cand.as<uint64_t>(),
                       w.n_chosen_s.as<uint32_t>(), w.rstart.as<uint32_t>(), w.rpre.as<uint32_t>(), S.lab.as<float>(),
                       w.ranges.as<float>(), w.asize.as<uint64_t>(), w.Lkeys.as<uint64_t>(), w.akeys.as<uint64_t>(),
                       w.lsize.as<uint64_t>(), w.L.as<uint32_t>(), w.nL.as<uint32_t>());
    HDG_HIP(hipGetLastError());
  }
  mark();
  hipLaunchKernelGGL(final_kernel, dim3((unsigned)qn), dim3(256), sort_lds, s, qn, q0, queries, d,
                     S.vec.as<float>(), S.zid.as<uint32_t>(), l, topk, w.mode.as<uint32_t>(), w.kprime.as<uint64_t>(),
                     w.lsize.as<uint64_t>(), w.L.as<uint32_t>(), w.nL.as<uint32_t>(), w.pad32.as<uint64_t>(),
                     w.padmax.as<uint64_t>(), w.Lkeys.as<uint64_t>(), w.akeys.as<uint64_t>(), w.asize.as<uint64_t>(),
                     out_dist, out_ids);
  HDG_HIP(hipGetLastError());
  mark();
  // per-query statistics (host copies)
  {
    std::vector<uint64_t> cand(qn), real(qn), kp(qn);
    std::vector<uint32_t> nl(qn), mode(qn), nc(qn), ns(qn);
    HDG_HIP(hipMemcpyAsync(cand.data(), w.cand.data(), qn * 8, hipMemcpyDeviceToHost, s));
    HDG_HIP(hipMemcpyAsync(real.data(), w.real.data(), qn * 8, hipMemcpyDeviceToHost, s));
    HDG_HIP(hipMemcpyAsync(kp.data(), w.kprime.data(), qn * 8, hipMemcpyDeviceToHost, s));
    HDG_HIP(hipMemcpyAsync(nl.data(), w.nL.data(), qn * 4, hipMemcpyDeviceToHost, s));
    HDG_HIP(hipMemcpyAsync(mode.data(), w.mode.data(), qn * 4, hipMemcpyDeviceToHost, s));
    HDG_HIP(hipMemcpyAsync(nc.data(), lazy ? w.n_chosen_c.data() : w.nnz_c.data(), qn * 4, hipMemcpyDeviceToHost,
                           s));
    HDG_HIP(hipMemcpyAsync(ns.data(), w.nnz_s.data(), qn * 4, hipMemcpyDeviceToHost, s));
    stats_.chosen_subs.resize(qn * sub_lists);
    HDG_HIP(hipMemcpyAsync(stats_.chosen_subs.data(), w.chosen_s.data(), qn * sub_lists * 4, hipMemcpyDeviceToHost, s));
    HDG_HIP(hipStreamSynchronize(s));
    for (uint64_t i = 0; i < qn; ++i) {
      stats_.cand[q0 + i] = cand[i];
      stats_.real[q0 + i] = real[i];
      stats_.kprime[q0 + i] = kp[i];
      stats_.refined[q0 + i] = nl[i];
      stats_.adc[q0 + i] = mode[i] == 1 ? cand[i] : 0;
      stats_.scored[q0 + i] = (uint64_t)nc[i] + ns[i];
    }
    static const char* plain_names[] = {"prep_kernel",    "count_kernel",      "cluster_nnz_kernel",
                                        "coarse_select_kernel", "sub_nnz_kernel", "sub_select_kernel",
                                        "ranges_kernel",  "plan+sync",         "keep_all_kernel",
                                        "adc_kernels",    "final_kernel"};
    static const char* lazy_names[] = {"prep_kernel",    "coarse_gemm_kernel", "coarse_topk+budget",
                                       "budget_count_kernel", "sub_nnz_kernel", "sub_select_kernel",
                                       "ranges_kernel",  "plan+sync",          "keep_all_kernel",
                                       "adc_kernels",    "final_kernel"};
    static const char* mixed_names[] = {"prep_kernel",    "count_kernel",        "cluster_nnz_kernel",
                                        "coarse_gemm+select+budget", "sub_nnz_kernel", "sub_gemm+select",
                                        "ranges_kernel",  "plan+sync",           "keep_all_kernel",
                                        "adc_kernels",    "final_kernel"};
    const char* const* names = lazy ? lazy_names : (mixed ? mixed_names : plain_names);
    const int nnames = 11;
    for (int i = 0; i + 1 < evn && i < nnames; ++i) {
      float ms = 0;
      HDG_HIP(hipEventElapsedTime(&ms, w.ev[i], w.ev[i + 1]));
      stage_ms.push_back({names[i], ms});
    }
  }
}

}  // namespace hdg
