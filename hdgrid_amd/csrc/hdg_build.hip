// Index build: Super-Grid-IVF_PQ.cu:420-646 (index::build) and product-quantization.cu:147-212
// (build_pq), with the RAFT k-means calls replaced by this file's own k-means:
//   * assignment on the GPU (tiled FP32 ||c||^2 - 2 x.c, ties to the lower center id);
//   * centroid update accumulated in FP64 on the GPU, divided on the host;
//   * empty clusters re-seeded from random training rows.
// The deterministic parts (grid shape, cells, Morton codes, cluster-major ordering, per-sub-cluster
// Morton sort) follow the reference exactly; the k-means outcome is not pinned by the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>

#include "hdg_index.hpp"

namespace hdg {

namespace {

constexpr int kTile = 64;   // rows / centers per tile
constexpr int kChunk = 32;  // dims per LDS stage

// Nearest center for each row of X (rows of width w, row stride ld).  256 threads: thread (ty,tx)
// owns rows ty*4..+3 and centers tx*4..+3 of the current 64x64 tile.
__global__ __launch_bounds__(256) void assign_kernel(const float* __restrict__ X, uint64_t n,
                                                     uint64_t ld, uint32_t w,
                                                     const float* __restrict__ Cn, uint32_t k,
                                                     const float* __restrict__ cnorm,
                                                     uint32_t* __restrict__ out_label,
                                                     float* __restrict__ out_score) {
  __shared__ float xs[kChunk][kTile + 1];
  __shared__ float cs[kChunk][kTile + 1];
  __shared__ float red_v[16][kTile];
  __shared__ uint32_t red_i[16][kTile];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const uint64_t row0 = (uint64_t)blockIdx.x * kTile;
  float best_v[4];
  uint32_t best_i[4];
  for (int r = 0; r < 4; ++r) {
    best_v[r] = INFINITY;
    best_i[r] = 0xFFFFFFFFu;
  }
  for (uint32_t c0 = 0; c0 < k; c0 += kTile) {
    float acc[4][4] = {};
    for (uint32_t t0 = 0; t0 < w; t0 += kChunk) {
      // stage 64 rows x 32 dims of X and of the centers (zero padded)
      for (int e = threadIdx.x; e < kTile * kChunk; e += 256) {
        int r = e / kChunk, t = e % kChunk;
        uint64_t row = row0 + r;
        uint32_t tt = t0 + t;
        xs[t][r] = (row < n && tt < w) ? X[row * ld + tt] : 0.0f;
        uint32_t c = c0 + r;
        cs[t][r] = (c < k && tt < w) ? Cn[(uint64_t)c * w + tt] : 0.0f;
      }
      __syncthreads();
#pragma unroll 8
      for (int t = 0; t < kChunk; ++t) {
        float xv[4], cv[4];
        for (int r = 0; r < 4; ++r) xv[r] = xs[t][ty * 4 + r];
        for (int c = 0; c < 4; ++c) cv[c] = cs[t][tx * 4 + c];
        for (int r = 0; r < 4; ++r)
          for (int c = 0; c < 4; ++c) acc[r][c] = fmaf(xv[r], cv[c], acc[r][c]);
      }
      __syncthreads();
    }
    for (int c = 0; c < 4; ++c) {
      uint32_t cid = c0 + tx * 4 + c;
      if (cid >= k) continue;
      float cn = cnorm[cid];
      for (int r = 0; r < 4; ++r) {
        float s = cn - 2.0f * acc[r][c];
        if (s < best_v[r] || (s == best_v[r] && cid < best_i[r])) {
          best_v[r] = s;
          best_i[r] = cid;
        }
      }
    }
  }
  for (int r = 0; r < 4; ++r) {
    red_v[tx][ty * 4 + r] = best_v[r];
    red_i[tx][ty * 4 + r] = best_i[r];
  }
  __syncthreads();
  if (threadIdx.x < kTile) {
    int r = threadIdx.x;
    float bv = red_v[0][r];
    uint32_t bi = red_i[0][r];
    for (int j = 1; j < 16; ++j) {
      float v = red_v[j][r];
      uint32_t i = red_i[j][r];
      if (v < bv || (v == bv && i < bi)) {
        bv = v;
        bi = i;
      }
    }
    uint64_t row = row0 + r;
    if (row < n) {
      out_label[row] = bi;
      if (out_score) out_score[row] = bv;
    }
  }
}

// FP64 centroid sums: sums[label][t] += X[row][t]; counts[label] += 1.
__global__ void accumulate_kernel(const float* __restrict__ X, uint64_t n, uint64_t ld, uint32_t w,
                                  const uint32_t* __restrict__ label, double* __restrict__ sums,
                                  unsigned long long* __restrict__ counts) {
  uint64_t row = (uint64_t)blockIdx.x * blockDim.y + threadIdx.y;
  if (row >= n) return;
  uint32_t c = label[row];
  for (uint32_t t = threadIdx.x; t < w; t += blockDim.x)
    atomicAdd(&sums[(uint64_t)c * w + t], (double)X[row * ld + t]);
  if (threadIdx.x == 0) atomicAdd(&counts[c], 1ull);
}

// dst[i] = src[idx[i]] for rows of `words` u32 words.
__global__ void gather_rows_kernel(const uint32_t* __restrict__ src, uint64_t src_ld_words,
                                   const uint32_t* __restrict__ idx, uint64_t n, uint32_t words,
                                   uint32_t* __restrict__ dst) {
  uint64_t row = (uint64_t)blockIdx.x * blockDim.y + threadIdx.y;
  if (row >= n) return;
  const uint32_t* s = src + (uint64_t)idx[row] * src_ld_words;
  uint32_t* o = dst + row * words;
  for (uint32_t t = threadIdx.x; t < words; t += blockDim.x) o[t] = s[t];
}

// Strided sub-matrix copy: dst[i][0..w) = src[i*ld + off + 0..w).
__global__ void slice_kernel(const float* __restrict__ src, uint64_t n, uint64_t ld, uint64_t off,
                             uint32_t w, float* __restrict__ dst) {
  uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * w) return;
  uint64_t r = e / w, t = e % w;
  dst[e] = src[r * ld + off + t];
}

uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

void launch_assign(const float* X, uint64_t n, uint64_t ld, uint32_t w, const float* Cn, uint32_t k,
                   const float* cnorm, uint32_t* out_label, float* out_score, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(assign_kernel, dim3((unsigned)ceil_div(n, kTile)), dim3(256), 0, s, X, n, ld, w,
                     Cn, k, cnorm, out_label, out_score);
  HDG_HIP(hipGetLastError());
}

// Lloyd k-means on X (device, rows n, width w, stride ld).  centers (host, [k][w]) are seeded here
// from distinct random rows.  Returns the final centers on the host.
void kmeans(const float* X, uint64_t n, uint64_t ld, uint32_t w, uint32_t k, int iters, uint64_t seed,
            std::vector<float>& centers, hipStream_t s) {
  if (n < k) fail(kInvalidArgument, "k-means needs at least as many rows as centers");
  std::mt19937_64 rng(seed);
  centers.assign((size_t)k * w, 0.0f);
  // distinct random seed rows
  std::vector<uint64_t> pick;
  {
    std::vector<uint64_t> perm;
    if (n <= 4ull * k) {
      perm.resize(n);
      std::iota(perm.begin(), perm.end(), 0);
      std::shuffle(perm.begin(), perm.end(), rng);
      pick.assign(perm.begin(), perm.begin() + k);
    } else {
      std::vector<char> used;
      std::uniform_int_distribution<uint64_t> U(0, n - 1);
      std::vector<uint64_t> tmp;
      tmp.reserve(k);
      std::vector<uint64_t> sorted;
      while (tmp.size() < k) {
        uint64_t r = U(rng);
        auto it = std::lower_bound(sorted.begin(), sorted.end(), r);
        if (it != sorted.end() && *it == r) continue;
        sorted.insert(it, r);
        tmp.push_back(r);
      }
      pick = tmp;
    }
  }
  auto fetch_row = [&](uint64_t r, float* dst) {
    HDG_HIP(hipMemcpyAsync(dst, X + r * ld, w * sizeof(float), hipMemcpyDeviceToHost, s));
  };
  for (uint32_t c = 0; c < k; ++c) fetch_row(pick[c], &centers[(size_t)c * w]);
  HDG_HIP(hipStreamSynchronize(s));

  DevBuf d_c, d_norm, d_label, d_sums, d_cnt;
  d_c.reserve((size_t)k * w * sizeof(float));
  d_norm.reserve(k * sizeof(float));
  d_label.reserve(n * sizeof(uint32_t));
  d_sums.reserve((size_t)k * w * sizeof(double));
  d_cnt.reserve(k * sizeof(unsigned long long));
  std::vector<float> norm(k);
  std::vector<double> sums((size_t)k * w);
  std::vector<unsigned long long> cnt(k);
  std::uniform_int_distribution<uint64_t> U(0, n - 1);
  for (int it = 0; it < iters; ++it) {
    for (uint32_t c = 0; c < k; ++c) {
      double a = 0;
      for (uint32_t t = 0; t < w; ++t) a += (double)centers[(size_t)c * w + t] * centers[(size_t)c * w + t];
      norm[c] = (float)a;
    }
    HDG_HIP(hipMemcpyAsync(d_c.data(), centers.data(), (size_t)k * w * sizeof(float), hipMemcpyHostToDevice, s));
    HDG_HIP(hipMemcpyAsync(d_norm.data(), norm.data(), k * sizeof(float), hipMemcpyHostToDevice, s));
    launch_assign(X, n, ld, w, d_c.as<float>(), k, d_norm.as<float>(), d_label.as<uint32_t>(), nullptr, s);
    HDG_HIP(hipMemsetAsync(d_sums.data(), 0, (size_t)k * w * sizeof(double), s));
    HDG_HIP(hipMemsetAsync(d_cnt.data(), 0, k * sizeof(unsigned long long), s));
    dim3 blk(32, 8);
    hipLaunchKernelGGL(accumulate_kernel, dim3((unsigned)ceil_div(n, 8)), blk, 0, s, X, n, ld, w,
                       d_label.as<uint32_t>(), d_sums.as<double>(), d_cnt.as<unsigned long long>());
    HDG_HIP(hipGetLastError());
    HDG_HIP(hipMemcpyAsync(sums.data(), d_sums.data(), (size_t)k * w * sizeof(double), hipMemcpyDeviceToHost, s));
    HDG_HIP(hipMemcpyAsync(cnt.data(), d_cnt.data(), k * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HDG_HIP(hipStreamSynchronize(s));
    for (uint32_t c = 0; c < k; ++c) {
      if (cnt[c] == 0) {
        fetch_row(U(rng), &centers[(size_t)c * w]);  // re-seed an empty cluster
        continue;
      }
      for (uint32_t t = 0; t < w; ++t)
        centers[(size_t)c * w + t] = (float)(sums[(size_t)c * w + t] / (double)cnt[c]);
    }
    HDG_HIP(hipStreamSynchronize(s));
  }
}

// Assign every row of X to its nearest center (centers on host).
void assign_all(const float* X, uint64_t n, uint64_t ld, uint32_t w, const std::vector<float>& centers,
                uint32_t k, DevBuf& d_label, hipStream_t s) {
  std::vector<float> norm(k);
  for (uint32_t c = 0; c < k; ++c) {
    double a = 0;
    for (uint32_t t = 0; t < w; ++t) a += (double)centers[(size_t)c * w + t] * centers[(size_t)c * w + t];
    norm[c] = (float)a;
  }
  DevBuf d_c, d_norm;
  upload(d_c, centers, s);
  upload(d_norm, norm, s);
  d_label.reserve(n * sizeof(uint32_t));
  launch_assign(X, n, ld, w, d_c.as<float>(), k, d_norm.as<float>(), d_label.as<uint32_t>(), nullptr, s);
  HDG_HIP(hipStreamSynchronize(s));
}

// Gather `rows` (host list) of X into a contiguous device matrix.
void gather_rows_dev(const float* X, uint64_t ld_floats, const std::vector<uint32_t>& rows, uint32_t w,
                     DevBuf& out, hipStream_t s) {
  DevBuf d_idx;
  upload(d_idx, rows, s);
  out.reserve(rows.size() * w * sizeof(float));
  if (rows.empty()) return;
  dim3 blk(64, 4);
  hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)ceil_div(rows.size(), 4)), blk, 0, s,
                     reinterpret_cast<const uint32_t*>(X), ld_floats, d_idx.as<uint32_t>(),
                     (uint64_t)rows.size(), w, out.as<uint32_t>());
  HDG_HIP(hipGetLastError());
  HDG_HIP(hipStreamSynchronize(s));
}

std::vector<uint32_t> sample_rows(uint64_t n, uint64_t want, uint64_t seed) {
  std::vector<uint32_t> rows;
  if (want >= n) {
    rows.resize(n);
    std::iota(rows.begin(), rows.end(), 0u);
    return rows;
  }
  // deterministic sample: reservoir-free via a shuffled stride permutation
  std::mt19937_64 rng(seed);
  std::vector<uint32_t> all(n);
  std::iota(all.begin(), all.end(), 0u);
  for (uint64_t i = 0; i < want; ++i) {
    std::uniform_int_distribution<uint64_t> U(i, n - 1);
    std::swap(all[i], all[U(rng)]);
  }
  rows.assign(all.begin(), all.begin() + want);
  std::sort(rows.begin(), rows.end());
  return rows;
}

}  // namespace

void build_bucket_grid(IndexState& st, const std::vector<uint32_t>& h_cells, hipStream_t s) {
  const GridShape& g = st.grid;
  const int nv = g.n_valid;
  const uint64_t N = st.N;
  st.Nb = 0;
  st.sat_per = 0;
  if (nv == 0 || N == 0) return;
  // about four points per bucket; the share is handed out from the dim with the fewest cells up, so a
  // dim with few cells keeps one bucket per cell and leaves the rest to the wide dims
  std::vector<int> order(nv);
  std::iota(order.begin(), order.end(), 0);
  for (int j = 0; j < nv; ++j) st.Mv[j] = (uint32_t)g.dims[g.valid_dims[j]];
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return st.Mv[a] < st.Mv[b]; });
  double remain = std::max(1.0, (double)N / 4.0);
  for (int r = 0; r < nv; ++r) {
    const int j = order[r];
    const double share = std::pow(remain, 1.0 / (double)(nv - r));
    st.nb[j] = (uint32_t)std::max(1.0, std::min<double>(st.Mv[j], std::floor(share + 0.5)));
    remain = std::max(1.0, remain / st.nb[j]);
  }
  st.bstride[nv - 1] = 1;
  for (int j = nv - 2; j >= 0; --j) st.bstride[j] = st.bstride[j + 1] * st.nb[j + 1];
  st.Nb = st.bstride[0] * st.nb[0];
  if (st.Nb >= (1ull << 32)) fail(kUnsupported, "bucket grid too large");
  // stable counting sort of z-pool positions by bucket
  std::vector<uint32_t> start(st.Nb + 1, 0), bid(N);
  for (uint64_t p = 0; p < N; ++p) {
    uint64_t b = 0;
    for (int j = 0; j < nv; ++j)
      b += ((uint64_t)h_cells[(uint64_t)j * N + p] * st.nb[j] / st.Mv[j]) * st.bstride[j];
    bid[p] = (uint32_t)b;
    start[b + 1]++;
  }
  for (uint64_t b = 0; b < st.Nb; ++b) start[b + 1] += start[b];
  std::vector<uint32_t> cur(start.begin(), start.end() - 1), ecell((uint64_t)nv * N), esub(N), epos(N);
  uint64_t sub = 0;
  for (uint64_t p = 0; p < N; ++p) {
    while (st.h_seg_off[sub + 1] <= p) ++sub;
    const uint32_t e = cur[bid[p]]++;
    for (int j = 0; j < nv; ++j) ecell[(uint64_t)j * N + e] = h_cells[(uint64_t)j * N + p];
    esub[e] = (uint32_t)sub;
    epos[e] = (uint32_t)p;
  }
  upload(st.bstart, start, s);
  upload(st.epos, epos, s);
  upload(st.ecell, ecell, s);
  upload(st.esub, esub, s);
  // coarse occupancy sums: 2^(12 / n_valid) coarse cells per dim (at most the dim's cells), counted per
  // sub-cluster, then prefix-summed along each dim
  const int gb = std::max(1, 12 / nv);
  for (int j = 0; j < nv; ++j) st.sg[j] = std::min<uint32_t>(1u << gb, st.Mv[j]);
  st.sstride[nv - 1] = 1;
  for (int j = nv - 2; j >= 0; --j) st.sstride[j] = st.sstride[j + 1] * (st.sg[j + 1] + 1);
  const uint64_t per = st.sstride[0] * (st.sg[0] + 1);
  if (per * st.CS > (1ull << 26)) return;  // more than 256 MiB: the dense count scans without it
  std::vector<uint32_t> t(per * st.CS, 0);
  sub = 0;
  for (uint64_t p = 0; p < N; ++p) {
    while (st.h_seg_off[sub + 1] <= p) ++sub;
    uint64_t x = 0;
    for (int j = 0; j < nv; ++j)
      x += ((uint64_t)h_cells[(uint64_t)j * N + p] * st.sg[j] / st.Mv[j] + 1) * st.sstride[j];
    t[sub * per + x]++;
  }
  for (int j = 0; j < nv; ++j) {
    const uint64_t w = st.sg[j] + 1, sj = st.sstride[j];
    for (uint64_t c = 0; c < st.CS; ++c) {
      uint32_t* tc = t.data() + c * per;
      for (uint64_t x = 0; x < per; ++x)
        if ((x / sj) % w != 0) tc[x] += tc[x - sj];
    }
  }
  upload(st.sat, t, s);
  st.sat_per = per;
}

// rows [n][d] f32 -> bf16 [n][Dp] (round to nearest even, zero padded) and |row_bf16|^2 in f32.
static void to_bf16_rows(const std::vector<float>& rows, uint64_t n, uint64_t d, uint64_t Dp, DevBuf& b16_out,
                         DevBuf& norm_out, hipStream_t s) {
  std::vector<uint16_t> b16(n * Dp, 0);
  std::vector<float> norm(n, 0.0f);
  for (uint64_t c = 0; c < n; ++c) {
    double acc = 0.0;
    for (uint64_t t = 0; t < d; ++t) {
      uint32_t u;
      float v = rows[c * d + t];
      std::memcpy(&u, &v, 4);
      u += 0x7FFFu + ((u >> 16) & 1u);
      const uint16_t h = (uint16_t)(u >> 16);
      b16[c * Dp + t] = h;
      uint32_t back = (uint32_t)h << 16;
      float r;
      std::memcpy(&r, &back, 4);
      acc += (double)r * r;
    }
    norm[c] = (float)acc;
  }
  upload(b16_out, b16, s);
  upload(norm_out, norm, s);
}

void prepare_mixed_centers(IndexState& st, const std::vector<float>& centers, const std::vector<float>& sub_centers,
                           hipStream_t s) {
  st.Dp = (st.d + 31) / 32 * 32;
  to_bf16_rows(centers, st.C, st.d, st.Dp, st.centers_bf16, st.center_norm, s);
  to_bf16_rows(sub_centers, st.CS, st.d, st.Dp, st.sub_centers_bf16, st.sub_center_norm, s);
}

void Index::build(const float* dataset, uint64_t n, uint64_t d, const float* labels, uint64_t label_cols,
                  hipStream_t s) {
  HDG_HIP(hipSetDevice(device_));
  drop_graphs();
  if (!dataset || !labels) fail(kInvalidArgument, "null dataset or labels");
  if (n == 0 || d == 0) fail(kInvalidArgument, "empty dataset");
  if (n >= 0xFFFFFFFFull) fail(kUnsupported, "more than 2^32-1 base vectors per index");
  if ((int)label_cols != fc_.l) fail(kInvalidArgument, "label columns differ from filter.conf l");
  const uint64_t C = cfg_.clusters, S = cfg_.sub_clusters, CS = C * S;
  const uint64_t M = cfg_.pq_dim, K = cfg_.n_clusters;
  if (C == 0 || S == 0) fail(kConfigError, "clusters and sub_clusters must be positive");
  if (M == 0 || K == 0 || K > 256) fail(kConfigError, "pq_dim must be positive and n_clusters <= 256");
  if (CS > n) fail(kInvalidArgument, "clusters*sub_clusters exceeds the number of base vectors");
  if (K > n) fail(kInvalidArgument, "n_clusters exceeds the number of base vectors");
  const uint64_t pq_len = (d + M - 1) / M;  // product-quantization.cu:161
  const int l = fc_.l;
  st_ = IndexState();
  st_.N = n;
  st_.d = d;
  st_.C = C;
  st_.CS = CS;
  st_.M = M;
  st_.K = K;
  st_.pq_len = pq_len;
  st_.l = l;

  // ---- B1: grid shape from the labels (host copy) ----
  std::vector<float> h_lab(n * l);
  HDG_HIP(hipMemcpyAsync(h_lab.data(), labels, n * l * sizeof(float), hipMemcpyDeviceToHost, s));
  HDG_HIP(hipStreamSynchronize(s));
  st_.grid = compute_grid_shape(h_lab.data(), n, l, cfg_.max_grids, cfg_.label_mask);
  const GridShape& g = st_.grid;

  // ---- B6a: PQ codebooks per subspace (raw vectors, product-quantization.cu:182-211) ----
  std::vector<float> pq_centers(M * K * pq_len, 0.0f);
  DevBuf d_codes_orig;  // u8 [n][M] original order
  d_codes_orig.reserve(n * M);
  {
    std::vector<uint32_t> rows = sample_rows(n, std::max<uint64_t>(cfg_.pq_train, K), cfg_.seed ^ 0x5151);
    DevBuf d_sub, d_label;
    d_sub.reserve(std::max<uint64_t>(rows.size(), n) * pq_len * sizeof(float));
    DevBuf d_train;
    gather_rows_dev(dataset, d, rows, (uint32_t)d, d_train, s);
    std::vector<uint8_t> codes_col;  // temporary
    DevBuf d_lab32;
    std::vector<uint32_t> h_l32(n);
    std::vector<uint8_t> h_codes(n * M);
    for (uint64_t m = 0; m < M; ++m) {
      const uint64_t off = m * pq_len;
      const uint32_t w = (uint32_t)std::min<uint64_t>(pq_len, d - std::min(d, off));
      std::vector<float> cent;
      if (w == 0) {
        cent.assign(K * pq_len, 0.0f);
      } else {
        // training slice [rows][w]
        hipLaunchKernelGGL(slice_kernel, dim3((unsigned)ceil_div(rows.size() * w, 256)), dim3(256), 0, s,
                           d_train.as<float>(), (uint64_t)rows.size(), d, off, w, d_sub.as<float>());
        HDG_HIP(hipGetLastError());
        std::vector<float> cw;
        kmeans(d_sub.as<float>(), rows.size(), w, w, (uint32_t)K, (int)cfg_.kmeans_iters,
               cfg_.seed * 1315423911ull + m, cw, s);
        cent.assign(K * pq_len, 0.0f);
        for (uint64_t c = 0; c < K; ++c)
          for (uint32_t t = 0; t < w; ++t) cent[c * pq_len + t] = cw[c * w + t];
        // encode all n rows of this subspace
        hipLaunchKernelGGL(slice_kernel, dim3((unsigned)ceil_div(n * w, 256)), dim3(256), 0, s, dataset, n, d,
                           off, w, d_sub.as<float>());
        HDG_HIP(hipGetLastError());
        assign_all(d_sub.as<float>(), n, w, w, cw, (uint32_t)K, d_lab32, s);
        HDG_HIP(hipMemcpy(h_l32.data(), d_lab32.data(), n * sizeof(uint32_t), hipMemcpyDeviceToHost));
      }
      for (uint64_t p = 0; p < n; ++p) h_codes[p * M + m] = w == 0 ? 0 : (uint8_t)h_l32[p];
      std::copy(cent.begin(), cent.end(), pq_centers.begin() + m * K * pq_len);
    }
    upload(d_codes_orig, h_codes, s);
  }

  // ---- B6b: fine k-means (CS centers), Super-Grid-IVF_PQ.cu:431-463 ----
  std::vector<float> fine;
  DevBuf d_fine_label;
  {
    std::vector<uint32_t> rows = sample_rows(n, std::max<uint64_t>(cfg_.kmeans_train, CS), cfg_.seed ^ 0xF1E);
    DevBuf d_train;
    gather_rows_dev(dataset, d, rows, (uint32_t)d, d_train, s);
    kmeans(d_train.as<float>(), rows.size(), d, (uint32_t)d, (uint32_t)CS, (int)cfg_.kmeans_iters,
           cfg_.seed + 7, fine, s);
    assign_all(dataset, n, d, (uint32_t)d, fine, (uint32_t)CS, d_fine_label, s);
  }
  std::vector<uint32_t> fine_label(n);
  HDG_HIP(hipMemcpy(fine_label.data(), d_fine_label.data(), n * sizeof(uint32_t), hipMemcpyDeviceToHost));

  // ---- B6c: group sub-centers into C clusters (stands in for kmeans_balanced, :466-476) ----
  std::vector<float> coarse;
  std::vector<uint32_t> group(CS);
  {
    DevBuf d_fine;
    upload(d_fine, fine, s);
    kmeans(d_fine.as<float>(), CS, d, (uint32_t)d, (uint32_t)C, (int)cfg_.balance_iters, cfg_.seed + 11, coarse, s);
    DevBuf d_g;
    assign_all(d_fine.as<float>(), CS, d, (uint32_t)d, coarse, (uint32_t)C, d_g, s);
    HDG_HIP(hipMemcpy(group.data(), d_g.data(), CS * sizeof(uint32_t), hipMemcpyDeviceToHost));
  }

  // ---- cluster-major sub-cluster order (group_by_cluster_id keeps ascending ids), :481-504 ----
  std::vector<uint32_t> sub_order;  // position -> fine id
  st_.h_sub_off.assign(C + 1, 0);
  {
    std::vector<std::vector<uint32_t>> members(C);
    for (uint32_t f = 0; f < CS; ++f) members[group[f]].push_back(f);
    for (uint64_t c = 0; c < C; ++c) {
      st_.h_sub_off[c] = (uint32_t)sub_order.size();
      for (uint32_t f : members[c]) sub_order.push_back(f);
      st_.max_subs_per_cluster = std::max<int>(st_.max_subs_per_cluster, (int)members[c].size());
    }
    st_.h_sub_off[C] = (uint32_t)sub_order.size();
  }
  std::vector<uint32_t> pos_of_fine(CS);
  for (uint32_t p = 0; p < CS; ++p) pos_of_fine[sub_order[p]] = p;
  std::vector<float> sub_centers(CS * d);
  for (uint64_t p = 0; p < CS; ++p)
    std::copy(fine.begin() + (size_t)sub_order[p] * d, fine.begin() + (size_t)(sub_order[p] + 1) * d,
              sub_centers.begin() + p * d);
  std::vector<uint32_t> cluster_of_sub(CS);
  for (uint64_t c = 0; c < C; ++c)
    for (uint32_t p = st_.h_sub_off[c]; p < st_.h_sub_off[c + 1]; ++p) cluster_of_sub[p] = (uint32_t)c;

  // ---- B2/B5: cells and Morton codes; z-pool = per sub-cluster ascending (zcode, id) ----
  std::vector<int> coords(n * l);
  std::vector<uint64_t> zc(n);
  for (uint64_t p = 0; p < n; ++p) {
    int* cp = &coords[p * l];
    for (int i = 0; i < l; ++i) {
      int gcell = grid_cell(h_lab[p * l + g.shuffle[i]], g.left[i], g.span[i], g.dims[i]);
      cp[i] = std::min(gcell, g.dims[i] - 1);
    }
    zc[p] = morton_code(cp, l, g.max_bits, g.valid_map);
  }
  st_.h_seg_off.assign(CS + 1, 0);
  {
    std::vector<uint64_t> cnt(CS, 0);
    for (uint64_t p = 0; p < n; ++p) cnt[pos_of_fine[fine_label[p]]]++;
    for (uint64_t s2 = 0; s2 < CS; ++s2) st_.h_seg_off[s2 + 1] = st_.h_seg_off[s2] + cnt[s2];
    st_.h_zid.assign(n, 0);
    std::vector<uint64_t> fillp(st_.h_seg_off.begin(), st_.h_seg_off.end() - 1);
    for (uint64_t p = 0; p < n; ++p) st_.h_zid[fillp[pos_of_fine[fine_label[p]]]++] = (uint32_t)p;
    for (uint64_t s2 = 0; s2 < CS; ++s2) {
      auto b = st_.h_zid.begin() + st_.h_seg_off[s2], e = st_.h_zid.begin() + st_.h_seg_off[s2 + 1];
      std::sort(b, e, [&](uint32_t a, uint32_t c) { return zc[a] != zc[c] ? zc[a] < zc[c] : a < c; });
      st_.max_seg_len = std::max<uint32_t>(st_.max_seg_len, (uint32_t)(e - b));
    }
    st_.h_zcode.resize(n);
    for (uint64_t p = 0; p < n; ++p) st_.h_zcode[p] = zc[st_.h_zid[p]];
  }

  // ---- device arrays in z-pool order ----
  std::vector<uint32_t> h_cells((uint64_t)g.n_valid * n);
  std::vector<float> h_zlab(n * l);
  for (uint64_t p = 0; p < n; ++p) {
    uint32_t id = st_.h_zid[p];
    for (int j = 0; j < g.n_valid; ++j) h_cells[(uint64_t)j * n + p] = (uint32_t)coords[(uint64_t)id * l + g.valid_dims[j]];
    for (int i = 0; i < l; ++i) h_zlab[p * l + i] = h_lab[(uint64_t)id * l + i];
  }
  upload(st_.cells, h_cells, s);
  build_bucket_grid(st_, h_cells, s);
  st_.labels_integral = std::all_of(h_lab.begin(), h_lab.end(), [](float v) { return v == std::floor(v); });
  upload(st_.lab, h_zlab, s);
  upload(st_.zcode, st_.h_zcode, s);
  upload(st_.zid, st_.h_zid, s);
  upload(st_.pq_centers, pq_centers, s);
  upload(st_.centers, coarse, s);
  prepare_mixed_centers(st_, coarse, sub_centers, s);
  upload(st_.sub_centers, sub_centers, s);
  upload(st_.sub_off, st_.h_sub_off, s);
  {
    std::vector<uint32_t> so(CS + 1);
    for (uint64_t i = 0; i <= CS; ++i) so[i] = (uint32_t)st_.h_seg_off[i];
    upload(st_.seg_off, so, s);
  }
  upload(st_.cluster_of_sub, cluster_of_sub, s);
  // vectors and PQ codes gathered into z-pool order on the device
  {
    DevBuf d_zid;
    upload(d_zid, st_.h_zid, s);
    st_.vec.reserve(n * d * sizeof(float));
    dim3 blk(64, 4);
    hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)ceil_div(n, 4)), blk, 0, s,
                       reinterpret_cast<const uint32_t*>(dataset), d, d_zid.as<uint32_t>(), n, (uint32_t)d,
                       st_.vec.as<uint32_t>());
    HDG_HIP(hipGetLastError());
    std::vector<uint8_t> h_codes(n * M), h_zcodes(n * M);
    HDG_HIP(hipMemcpyAsync(h_codes.data(), d_codes_orig.data(), n * M, hipMemcpyDeviceToHost, s));
    HDG_HIP(hipStreamSynchronize(s));
    for (uint64_t p = 0; p < n; ++p)
      std::memcpy(&h_zcodes[p * M], &h_codes[(uint64_t)st_.h_zid[p] * M], M);
    upload(st_.codes, h_zcodes, s);
  }
  prepare_blocked_codes(st_, s);
  HDG_HIP(hipStreamSynchronize(s));
  st_.built = true;
}

namespace {
// One thread per row: its sixteen-byte words and its eight-byte tail move to their places in the 64-row block.
__global__ void block_codes_kernel(const uint8_t* __restrict__ codes, uint64_t n, uint64_t M,
                                   uint8_t* __restrict__ out) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint64_t n16 = M / 16;
  const uint8_t* row = codes + p * M;  // rows are 8-byte aligned (M % 8 == 0)
  uint8_t* blk = out + (p >> 6) * 64 * M;
  for (uint64_t w = 0; w < n16; ++w) {
    const uint2 a = reinterpret_cast<const uint2*>(row)[2 * w], b = reinterpret_cast<const uint2*>(row)[2 * w + 1];
    reinterpret_cast<uint4*>(blk + w * 1024)[p & 63] = make_uint4(a.x, a.y, b.x, b.y);
  }
  if (M & 15) reinterpret_cast<uint2*>(blk + n16 * 1024)[p & 63] = reinterpret_cast<const uint2*>(row)[2 * n16];
}
}  // namespace

void prepare_blocked_codes(IndexState& st, hipStream_t s) {
  st.codes_wb = (st.M % 8) == 0 ? 16u : 0u;
  if (!st.codes_wb) return;
  const uint64_t rows = (st.N + 63) / 64 * 64;
  st.codes_blk.reserve(rows * st.M);
  hipLaunchKernelGGL(block_codes_kernel, dim3((unsigned)((st.N + 255) / 256)), dim3(256), 0, s,
                     st.codes.as<uint8_t>(), st.N, st.M, st.codes_blk.as<uint8_t>());
  HDG_HIP(hipGetLastError());
}

}  // namespace hdg
