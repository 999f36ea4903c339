// Index build: Super-Grid-IVF_PQ.cu:420-646 (index::build) and product-quantization.cu:147-212
// (build_pq), with the RAFT k-means calls replaced by this file's own k-means, all on the device:
//   * assignment: tiled FP32 ||c||^2 - 2 x.c, ties to the lower center id;
//   * centroid update: sums in 64-bit fixed point with integer atomics (associative, so independent of
//     the order rows arrive in), divided in FP64;
//   * empty clusters re-seeded from random training rows drawn on the host.
// Cells, Morton codes, the z-pool order (two stable radix sorts), the gathers into z-pool order, the
// bucket grid and the coarse occupancy sums run on the device too.  The deterministic parts (grid shape,
// cells, Morton codes, cluster-major ordering, per-sub-cluster Morton sort) follow the reference exactly;
// the k-means outcome is not pinned by the reference, but the same input and config give the same index
// state bit for bit.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cfloat>
#include <chrono>
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
  for (int r = 0; r < 4; ++r) {  // a row whose scores are all infinite or NaN keeps center 0
    best_v[r] = INFINITY;
    best_i[r] = 0;
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

// Largest |x| over the rows, as float bits (non-negative floats order like their bit patterns).
__global__ void absmax_kernel(const float* __restrict__ X, uint64_t n, uint64_t ld, uint32_t w,
                              unsigned int* __restrict__ out) {
  unsigned int m = 0;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * w; e += (uint64_t)gridDim.x * blockDim.x) {
    const float v = fabsf(X[(e / w) * ld + e % w]);
    const unsigned int b = v != v ? 0x7F800001u : __float_as_uint(v);  // NaN counts as non-finite
    m = b > m ? b : m;
  }
  atomicMax(out, m);
}

// Centroid sums in 64-bit fixed point: each coordinate is rounded once to a multiple of 1 / scale, where
// scale = 2^(62 - e) and 2^e > n * max|x|, so no partial sum leaves int64; the integer additions are
// associative, so sums[label][t] and counts[label] come out the same whatever order the rows arrive in.
__global__ void accumulate_kernel(const float* __restrict__ X, uint64_t n, uint64_t ld, uint32_t w,
                                  const uint32_t* __restrict__ label, double scale,
                                  unsigned long long* __restrict__ sums, unsigned long long* __restrict__ counts) {
  uint64_t row = (uint64_t)blockIdx.x * blockDim.y + threadIdx.y;
  if (row >= n) return;
  uint32_t c = label[row];
  for (uint32_t t = threadIdx.x; t < w; t += blockDim.x) {
    const long long v = __double2ll_rn((double)X[row * ld + t] * scale);
    atomicAdd(&sums[(uint64_t)c * w + t], (unsigned long long)v);  // two's complement: a signed sum
  }
  if (threadIdx.x == 0) atomicAdd(&counts[c], 1ull);
}

// centers[c][t] = sums[c][t] / scale / counts[c] for the non-empty clusters; an empty cluster keeps its
// center, or gets the zero vector when zero_empty is set (the balanced k-means' division rule).
__global__ void center_update_kernel(const unsigned long long* __restrict__ sums,
                                     const unsigned long long* __restrict__ counts, uint32_t k, uint32_t w,
                                     double inv_scale, float* __restrict__ centers, int zero_empty = 0) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (uint64_t)k * w) return;
  const unsigned long long cnt = counts[e / w];
  if (cnt == 0) {
    if (zero_empty) centers[e] = 0.0f;
    return;
  }
  centers[e] = (float)((double)(long long)sums[e] * inv_scale / (double)cnt);
}

// label[i] = i mod k: the balanced k-means' starting assignment.
__global__ void mod_label_kernel(uint32_t* __restrict__ label, uint64_t n, uint32_t k) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) label[i] = (uint32_t)(i % k);
}

// One balancing move per blockIdx.y: centers[l] = (wc * centers[li] + X[i]) / (wc + 1), each step rounded
// to f32.  The moved clusters are small and the sources large, so no block reads a center another writes.
__global__ void adjust_move_kernel(float* __restrict__ centers, uint32_t w, const uint32_t* __restrict__ moves,
                                   const float* __restrict__ wcs, const float* __restrict__ X, uint64_t ld) {
  const uint32_t m = blockIdx.y;
  const uint32_t l = moves[3 * m], li = moves[3 * m + 1], i = moves[3 * m + 2];
  const float wc = wcs[m];
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < w; t += gridDim.x * blockDim.x) {
    float v = __fmul_rn(wc, centers[(uint64_t)li * w + t]);
    v = __fadd_rn(v, X[(uint64_t)i * ld + t]);
    centers[(uint64_t)l * w + t] = __fdiv_rn(v, __fadd_rn(wc, 1.0f));
  }
}

// |c|^2 of each center, summed in FP64 over the dims in order (one thread per center).
__global__ void center_norm_kernel(const float* __restrict__ centers, uint32_t k, uint32_t w,
                                   float* __restrict__ norm) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k) return;
  double a = 0.0;
  for (uint32_t t = 0; t < w; ++t) {
    const double v = centers[(uint64_t)c * w + t];
    a += v * v;
  }
  norm[c] = (float)a;
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

// dst[idx[i]] = src[i] for rows of w floats.
__global__ void scatter_rows_kernel(const float* __restrict__ src, const uint32_t* __restrict__ idx, uint64_t n,
                                    uint32_t w, float* __restrict__ dst) {
  uint64_t row = (uint64_t)blockIdx.x * blockDim.y + threadIdx.y;
  if (row >= n) return;
  float* o = dst + (uint64_t)idx[row] * w;
  for (uint32_t t = threadIdx.x; t < w; t += blockDim.x) o[t] = src[row * w + t];
}

// Strided sub-matrix copy: dst[i][0..w) = src[i*ld + off + 0..w).
__global__ void slice_kernel(const float* __restrict__ src, uint64_t n, uint64_t ld, uint64_t off,
                             uint32_t w, float* __restrict__ dst) {
  uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * w) return;
  uint64_t r = e / w, t = e % w;
  dst[e] = src[r * ld + off + t];
}

// PQ codes of every row for all subspaces in one pass over the data (product-quantization.cu:182-211 encodes
// one subspace at a time).  Thread = row.  For each subspace the block stages the K centers [K][pq_len]
// (zero padded past the subspace's width) and their norms in LDS; the code is the center with the smallest
// |c|^2 - 2 x.c, x.c an FMA chain over the dims in order, ties to the lower id: assign_kernel's rule, so
// the codes equal a per-subspace assign_kernel pass.  PQL >= pq_len is the register width of x.
template <int PQL>
__global__ __launch_bounds__(256) void pq_encode_kernel(const float* __restrict__ X, uint64_t n, uint64_t d,
                                                        uint32_t M, uint32_t K, uint32_t pq_len,
                                                        const float* __restrict__ pqc,
                                                        const float* __restrict__ pqn,
                                                        uint8_t* __restrict__ codes) {
  extern __shared__ float sm[];  // [K * pq_len] centers, then [K] norms
  const uint64_t row = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t m = 0; m < M; ++m) {
    const uint64_t off = (uint64_t)m * pq_len;
    const uint32_t w = (uint32_t)(off < d ? (d - off < pq_len ? d - off : pq_len) : 0);
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < K * pq_len; e += blockDim.x) sm[e] = pqc[(uint64_t)m * K * pq_len + e];
    for (uint32_t e = threadIdx.x; e < K; e += blockDim.x) sm[K * pq_len + e] = pqn[(uint64_t)m * K + e];
    __syncthreads();
    if (row >= n) continue;
    if (w == 0) {
      codes[row * M + m] = 0;
      continue;
    }
    float x[PQL];
#pragma unroll
    for (int t = 0; t < PQL; ++t) x[t] = (uint32_t)t < w ? X[row * d + off + t] : 0.0f;
    float best = INFINITY;
    uint32_t bi = 0;
    for (uint32_t c = 0; c < K; ++c) {
      const float* cc = sm + c * pq_len;
      float acc = 0.0f;
#pragma unroll
      for (int t = 0; t < PQL; ++t)
        if ((uint32_t)t < pq_len) acc = fmaf(x[t], cc[t], acc);
      const float sc = sm[K * pq_len + c] - 2.0f * acc;
      if (sc < best) {
        best = sc;
        bi = c;
      }
    }
    codes[row * M + m] = (uint8_t)bi;
  }
}

// codes[row][m] = label[row] (the per-subspace encoder for wide subspaces).
__global__ void code_column_kernel(const uint32_t* __restrict__ label, uint64_t n, uint32_t M, uint32_t m,
                                   uint8_t* __restrict__ codes) {
  const uint64_t row = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row < n) codes[row * M + m] = (uint8_t)label[row];
}

// Grid arguments for the cell / Morton kernel (the build's GridShape, by value).
struct GridArgs {
  int l, max_bits, valid_map, n_valid;
  int dims[kMaxDims], shuffle[kMaxDims], valid_dims[kMaxDims];
  uint64_t left[kMaxDims], span[kMaxDims];
};

// B2/B5 per row (Super-Grid-IVF_PQ.cu:559-600): the cell of each sorted dim (get_grid_id clamped to the
// last cell), the Morton code over the valid dims, the valid dims' cells ([n_valid][n], original order),
// the row's sub-cluster table position and its id (the value the sorts carry).
__global__ void cells_morton_kernel(const float* __restrict__ labels, uint64_t n, GridArgs g,
                                    const uint32_t* __restrict__ fine_label, const uint32_t* __restrict__ pos_of_fine,
                                    uint32_t* __restrict__ cells, uint64_t* __restrict__ zc,
                                    uint32_t* __restrict__ seg, uint32_t* __restrict__ ids) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int cp[kMaxDims];
  for (int i = 0; i < g.l; ++i) {
    const int c = grid_cell(labels[p * g.l + g.shuffle[i]], g.left[i], g.span[i], g.dims[i]);
    cp[i] = c < g.dims[i] - 1 ? c : g.dims[i] - 1;
  }
  zc[p] = morton_code(cp, g.l, g.max_bits, g.valid_map);
  for (int j = 0; j < g.n_valid; ++j) cells[(uint64_t)j * n + p] = (uint32_t)cp[g.valid_dims[j]];
  seg[p] = pos_of_fine[fine_label[p]];
  ids[p] = (uint32_t)p;
}

// out[i] = in[idx[i]] for elements of type T.
template <class T>
__global__ void gather_kernel(const T* __restrict__ in, const uint32_t* __restrict__ idx, uint64_t n,
                              T* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[idx[i]];
}

// dst[i][j] = src[idx[i]][j] for rows of `bytes` bytes.
__global__ void gather_bytes_kernel(const uint8_t* __restrict__ src, const uint32_t* __restrict__ idx, uint64_t n,
                                    uint32_t bytes, uint8_t* __restrict__ dst) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * bytes) return;
  const uint64_t i = e / bytes, j = e % bytes;
  dst[e] = src[(uint64_t)idx[i] * bytes + j];
}

// start[b] = first i with sorted[i] >= b, for b in [0, nkeys] (sorted ascending, n entries).
__global__ void lower_bounds_kernel(const uint32_t* __restrict__ sorted, uint64_t n, uint64_t nkeys,
                                    uint32_t* __restrict__ start) {
  const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nkeys) return;
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) / 2;
    if (sorted[mid] < b) lo = mid + 1;
    else hi = mid;
  }
  start[b] = (uint32_t)lo;
}

// The sub-cluster holding z-pool position p: the last s with seg_off[s] <= p (seg_off ascending, [CS+1]).
__device__ inline uint32_t segment_of(const uint32_t* __restrict__ seg_off, uint64_t CS, uint32_t p) {
  uint64_t lo = 0, hi = CS + 1;  // first s with seg_off[s] > p, minus one
  while (lo < hi) {
    const uint64_t mid = (lo + hi) / 2;
    if (seg_off[mid] <= p) lo = mid + 1;
    else hi = mid;
  }
  return (uint32_t)(lo - 1);
}

// Bucket grid arguments (the IndexState's bucket fields, by value).
struct BucketArgs {
  int nv;
  uint32_t Mv[kMaxDims], nb[kMaxDims], sg[kMaxDims];
  uint64_t bstride[kMaxDims], sstride[kMaxDims];
};

// bid[p] = bucket of z-pool position p; pos[p] = p.
__global__ void bucket_id_kernel(const uint32_t* __restrict__ cells, uint64_t N, BucketArgs a,
                                 uint32_t* __restrict__ bid, uint32_t* __restrict__ pos) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  uint64_t b = 0;
  for (int j = 0; j < a.nv; ++j) b += ((uint64_t)cells[(uint64_t)j * N + p] * a.nb[j] / a.Mv[j]) * a.bstride[j];
  bid[p] = (uint32_t)b;
  pos[p] = (uint32_t)p;
}

// Entry e of the bucket grid: its cells (type CT) and its sub-cluster, from the z-pool position epos[e]; with
// shift >= 0 the sub-cluster and the position are packed into one word, sub << shift | (p - seg_off[sub]).
template <class CT>
__global__ void entry_fill_kernel(const uint32_t* __restrict__ cells, const uint32_t* __restrict__ seg_off,
                                  uint64_t CS, uint64_t N, int nv, const uint32_t* __restrict__ epos, int shift,
                                  CT* __restrict__ ecell, uint32_t* __restrict__ esub) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N) return;
  const uint32_t p = epos[e];
  for (int j = 0; j < nv; ++j) ecell[(uint64_t)j * N + e] = (CT)cells[(uint64_t)j * N + p];
  const uint32_t sub = segment_of(seg_off, CS, p);
  esub[e] = shift >= 0 ? (sub << shift) | (p - seg_off[sub]) : sub;
}

// Coarse occupancy counts: t[sub][x] += 1 at x = sum_j (coarse cell_j + 1) * sstride[j].
__global__ void sat_count_kernel(const uint32_t* __restrict__ cells, const uint32_t* __restrict__ seg_off,
                                 uint64_t CS, uint64_t N, BucketArgs a, uint64_t per, uint32_t* __restrict__ t) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  uint64_t x = 0;
  for (int j = 0; j < a.nv; ++j)
    x += ((uint64_t)cells[(uint64_t)j * N + p] * a.sg[j] / a.Mv[j] + 1) * a.sstride[j];
  atomicAdd(&t[(uint64_t)segment_of(seg_off, CS, (uint32_t)p) * per + x], 1u);
}

// Inclusive prefix sums along dim j of every sub-cluster's table: one thread per line of the dim.
__global__ void sat_prefix_kernel(uint32_t* __restrict__ t, uint64_t lines, uint64_t per, uint64_t wj,
                                  uint64_t sj) {
  const uint64_t L = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (L >= lines) return;
  const uint64_t per_line = per / wj;  // lines of one sub-cluster's table
  const uint64_t c = L / per_line, r = L % per_line;
  uint32_t* base = t + c * per + (r / sj) * (wj * sj) + r % sj;
  uint32_t acc = base[0];
  for (uint64_t k = 1; k < wj; ++k) {
    acc += base[k * sj];
    base[k * sj] = acc;
  }
}

// Stable radix sort of (key, value) pairs on bits [0, end_bit) of the key; end_bit 0 copies.
template <class K>
void sort_pairs(const K* kin, K* kout, const uint32_t* vin, uint32_t* vout, uint64_t n, int end_bit, hipStream_t s) {
  if (n == 0) return;
  if (end_bit <= 0) {
    HDG_HIP(hipMemcpyAsync(kout, kin, n * sizeof(K), hipMemcpyDeviceToDevice, s));
    HDG_HIP(hipMemcpyAsync(vout, vin, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    return;
  }
  size_t bytes = 0;
  HDG_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)end_bit, s));
  DevBuf tmp;
  tmp.reserve(std::max<size_t>(bytes, 256));
  HDG_HIP(rocprim::radix_sort_pairs(tmp.data(), bytes, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)end_bit, s));
  HDG_HIP(hipStreamSynchronize(s));  // tmp is released on return
}

// Bits needed for keys in [0, nkeys).
int key_bits(uint64_t nkeys) {
  int b = 0;
  while (b < 64 && (nkeys - 1) >> b) ++b;
  return nkeys <= 1 ? 0 : b;
}

uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

void launch_assign(const float* X, uint64_t n, uint64_t ld, uint32_t w, const float* Cn, uint32_t k,
                   const float* cnorm, uint32_t* out_label, float* out_score, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(assign_kernel, dim3((unsigned)ceil_div(n, kTile)), dim3(256), 0, s, X, n, ld, w,
                     Cn, k, cnorm, out_label, out_score);
  HDG_HIP(hipGetLastError());
}

void gather_rows(const float* X, uint64_t ld, const DevBuf& d_idx, uint64_t rows, uint32_t w, float* out,
                 hipStream_t s) {
  if (rows == 0) return;
  hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)ceil_div(rows, 4)), dim3(64, 4), 0, s,
                     reinterpret_cast<const uint32_t*>(X), ld, d_idx.as<uint32_t>(), rows, w,
                     reinterpret_cast<uint32_t*>(out));
  HDG_HIP(hipGetLastError());
}

// Lloyd k-means on X (device, rows n, width w, stride ld), seeded from distinct random rows; the centers
// stay on the device between iterations and come back to the host at the end.  Every step is
// order-independent (assignment ties go to the lower id, sums are fixed point), so the same input and seed
// give the same centers bit for bit.
void kmeans(const float* X, uint64_t n, uint64_t ld, uint32_t w, uint32_t k, int iters, uint64_t seed,
            std::vector<float>& centers, hipStream_t s) {
  if (n < k) fail(kInvalidArgument, "k-means needs at least as many rows as centers");
  std::mt19937_64 rng(seed);
  // distinct random seed rows
  std::vector<uint32_t> pick;
  if (n <= 4ull * k) {
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    std::shuffle(perm.begin(), perm.end(), rng);
    pick.assign(perm.begin(), perm.begin() + k);
  } else {
    std::uniform_int_distribution<uint64_t> U(0, n - 1);
    std::vector<uint64_t> sorted;
    pick.reserve(k);
    while (pick.size() < k) {
      uint64_t r = U(rng);
      auto it = std::lower_bound(sorted.begin(), sorted.end(), r);
      if (it != sorted.end() && *it == r) continue;
      sorted.insert(it, r);
      pick.push_back((uint32_t)r);
    }
  }
  DevBuf d_c, d_norm, d_label, d_sums, d_cnt, d_idx, d_max;
  d_c.reserve((size_t)k * w * sizeof(float));
  d_norm.reserve(k * sizeof(float));
  d_label.reserve(n * sizeof(uint32_t));
  d_sums.reserve((size_t)k * w * sizeof(unsigned long long));
  d_cnt.reserve(k * sizeof(unsigned long long));
  upload(d_idx, pick, s);
  gather_rows(X, ld, d_idx, k, w, d_c.as<float>(), s);

  // fixed-point scale from the largest coordinate
  upload(d_max, std::vector<unsigned int>{0u}, s);
  hipLaunchKernelGGL(absmax_kernel, dim3((unsigned)std::min<uint64_t>(4096, ceil_div(n * w, 256))), dim3(256), 0, s,
                     X, n, ld, w, d_max.as<unsigned int>());
  HDG_HIP(hipGetLastError());
  unsigned int maxbits = 0;
  HDG_HIP(hipMemcpyAsync(&maxbits, d_max.data(), sizeof(maxbits), hipMemcpyDeviceToHost, s));
  HDG_HIP(hipStreamSynchronize(s));
  float maxabs;
  std::memcpy(&maxabs, &maxbits, sizeof(maxabs));
  if (!std::isfinite(maxabs)) fail(kInvalidArgument, "k-means input holds a non-finite value");
  int e = 0;
  std::frexp((double)maxabs * (double)n, &e);  // n * max|x| < 2^e
  const double scale = maxabs > 0.0f ? std::ldexp(1.0, 62 - e) : 1.0;

  std::vector<unsigned long long> cnt(k);
  std::uniform_int_distribution<uint64_t> U(0, n - 1);
  for (int it = 0; it < iters; ++it) {
    hipLaunchKernelGGL(center_norm_kernel, dim3((unsigned)ceil_div(k, 256)), dim3(256), 0, s, d_c.as<float>(), k, w,
                       d_norm.as<float>());
    HDG_HIP(hipGetLastError());
    launch_assign(X, n, ld, w, d_c.as<float>(), k, d_norm.as<float>(), d_label.as<uint32_t>(), nullptr, s);
    HDG_HIP(hipMemsetAsync(d_sums.data(), 0, (size_t)k * w * sizeof(unsigned long long), s));
    HDG_HIP(hipMemsetAsync(d_cnt.data(), 0, k * sizeof(unsigned long long), s));
    hipLaunchKernelGGL(accumulate_kernel, dim3((unsigned)ceil_div(n, 8)), dim3(32, 8), 0, s, X, n, ld, w,
                       d_label.as<uint32_t>(), scale, d_sums.as<unsigned long long>(),
                       d_cnt.as<unsigned long long>());
    HDG_HIP(hipGetLastError());
    hipLaunchKernelGGL(center_update_kernel, dim3((unsigned)ceil_div((uint64_t)k * w, 256)), dim3(256), 0, s,
                       d_sums.as<unsigned long long>(), d_cnt.as<unsigned long long>(), k, w, 1.0 / scale,
                       d_c.as<float>());
    HDG_HIP(hipGetLastError());
    HDG_HIP(hipMemcpyAsync(cnt.data(), d_cnt.data(), k * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HDG_HIP(hipStreamSynchronize(s));
    // re-seed the empty clusters from random rows, in ascending cluster order
    std::vector<uint32_t> empty_c, empty_r;
    for (uint32_t c = 0; c < k; ++c)
      if (cnt[c] == 0) {
        empty_c.push_back(c);
        empty_r.push_back((uint32_t)U(rng));
      }
    if (!empty_c.empty()) {
      DevBuf d_ec, d_er, d_tmp;
      upload(d_er, empty_r, s);
      d_tmp.reserve(empty_r.size() * w * sizeof(float));
      gather_rows(X, ld, d_er, empty_r.size(), w, d_tmp.as<float>(), s);
      upload(d_ec, empty_c, s);  // centers[empty_c[i]] = row empty_r[i]
      hipLaunchKernelGGL(scatter_rows_kernel, dim3((unsigned)ceil_div(empty_c.size(), 4)), dim3(64, 4), 0, s,
                         d_tmp.as<float>(), d_ec.as<uint32_t>(), (uint64_t)empty_c.size(), w, d_c.as<float>());
      HDG_HIP(hipGetLastError());
      HDG_HIP(hipStreamSynchronize(s));
    }
  }
  centers.resize((size_t)k * w);
  HDG_HIP(hipMemcpyAsync(centers.data(), d_c.data(), (size_t)k * w * sizeof(float), hipMemcpyDeviceToHost, s));
  HDG_HIP(hipStreamSynchronize(s));
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

// ---------------------------------------------------------------------------------------------------------
// Balanced hierarchical k-means: RAFT's kmeans_balanced::fit_predict as Super-Grid-IVF_PQ.cu:466-476 calls it
// (L2, n_iters 200) to group the clusters*sub_clusters fine centers into `clusters` groups.  RAFT is not
// vendored (rapids-cmake fetches it, CMakeLists.txt:18-34); this follows its published algorithm:
//   fit = mesoclusters (round(sqrt(k)) of them) trained on all rows; fine cluster counts handed to the
//         mesoclusters in proportion to their sizes; each mesocluster's rows clustered into its share; then
//         max(n_iters / 10, 2) balancing EM iterations over all rows with all k centers;
//   predict = nearest center.
// Training a set of clusters starts from labels i mod k and their means, then runs balancing EM: before
// every E step but the first, clusters of at most floor(average * threshold) rows (average = n / k) are
// moved next to a row of a cluster of at least `average` rows (centers[l] = (w * centers[li] + x_i) / (w + 1),
// w = min(size, 7)); every `pullback`-th iteration with such a move adds an iteration.  The row is found by
// stepping a cursor through the rows with a prime stride that does not divide n (the cursor and the stride
// index carry over between calls within one fit, in RAFT across the whole process).
// Arithmetic as this file's k-means: assignment by ||c||^2 - 2 x.c (ties to the lower id), fixed-point sums
// divided in FP64 (RAFT sums in f32), an empty cluster's center set to zero (RAFT divides by a zero count to
// zero).  Deterministic: the same rows and settings give the same centers and labels bit for bit.
void gather_rows_dev(const float* X, uint64_t ld_floats, const std::vector<uint32_t>& rows, uint32_t w,
                     DevBuf& out, hipStream_t s);

struct BalancedFit {
  hipStream_t s;
  uint64_t cursor = 0;  // row cursor of the balancing moves
  uint32_t prime_j = 0;  // index of the current stride prime
  DevBuf d_norm, d_sums, d_cnt, d_moves, d_wc;
  std::vector<unsigned long long> cnt;
  std::vector<uint32_t> lab;
};

// Stride primes: the smallest prime >= 29 + 70 j for j = 0..39 (29 ... 2,767).
uint64_t stride_prime(uint32_t j) {
  uint64_t p = 29 + 70ull * j;
  for (;; ++p) {
    bool prime = p > 1;
    for (uint64_t q = 2; q * q <= p && prime; ++q) prime = p % q != 0;
    if (prime) return p;
  }
}
constexpr uint32_t kStridePrimes = 40;
constexpr float kAdjustWeight = 7.0f;

double fixed_point_scale(const float* X, uint64_t n, uint64_t ld, uint32_t w, hipStream_t s) {
  DevBuf d_max;
  upload(d_max, std::vector<unsigned int>{0u}, s);
  if (n > 0 && w > 0) {
    hipLaunchKernelGGL(absmax_kernel, dim3((unsigned)std::min<uint64_t>(4096, ceil_div(n * w, 256))), dim3(256), 0,
                       s, X, n, ld, w, d_max.as<unsigned int>());
    HDG_HIP(hipGetLastError());
  }
  unsigned int maxbits = 0;
  HDG_HIP(hipMemcpyAsync(&maxbits, d_max.data(), sizeof(maxbits), hipMemcpyDeviceToHost, s));
  HDG_HIP(hipStreamSynchronize(s));
  float maxabs;
  std::memcpy(&maxabs, &maxbits, sizeof(maxabs));
  if (!std::isfinite(maxabs)) fail(kInvalidArgument, "k-means input holds a non-finite value");
  int e = 0;
  std::frexp((double)maxabs * (double)n, &e);  // n * max|x| < 2^e
  return maxabs > 0.0f ? std::ldexp(1.0, 62 - e) : 1.0;
}

// E step: label = nearest center.
void bal_predict(BalancedFit& f, const float* X, uint64_t n, uint64_t ld, uint32_t w, uint32_t k,
                 const float* centers, uint32_t* label) {
  f.d_norm.reserve(k * sizeof(float));
  hipLaunchKernelGGL(center_norm_kernel, dim3((unsigned)ceil_div(k, 256)), dim3(256), 0, f.s, centers, k, w,
                     f.d_norm.as<float>());
  HDG_HIP(hipGetLastError());
  launch_assign(X, n, ld, w, centers, k, f.d_norm.as<float>(), label, nullptr, f.s);
}

// M step: centers = means of their rows (zero when empty), sizes in f.d_cnt.
void bal_centers(BalancedFit& f, const float* X, uint64_t n, uint64_t ld, uint32_t w, uint32_t k,
                 const uint32_t* label, double scale, float* centers) {
  f.d_sums.reserve((size_t)k * w * sizeof(unsigned long long));
  f.d_cnt.reserve(k * sizeof(unsigned long long));
  HDG_HIP(hipMemsetAsync(f.d_sums.data(), 0, (size_t)k * w * sizeof(unsigned long long), f.s));
  HDG_HIP(hipMemsetAsync(f.d_cnt.data(), 0, k * sizeof(unsigned long long), f.s));
  if (n > 0) {
    hipLaunchKernelGGL(accumulate_kernel, dim3((unsigned)ceil_div(n, 8)), dim3(32, 8), 0, f.s, X, n, ld, w, label,
                       scale, f.d_sums.as<unsigned long long>(), f.d_cnt.as<unsigned long long>());
    HDG_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(center_update_kernel, dim3((unsigned)ceil_div((uint64_t)k * w, 256)), dim3(256), 0, f.s,
                     f.d_sums.as<unsigned long long>(), f.d_cnt.as<unsigned long long>(), k, w, 1.0 / scale,
                     centers, 1);
  HDG_HIP(hipGetLastError());
}

// Balancing move of every small cluster; true when a center moved.
bool bal_adjust(BalancedFit& f, const float* X, uint64_t n, uint64_t ld, uint32_t w, uint32_t k,
                const uint32_t* label, float threshold, float* centers) {
  if (k == 0 || n == 0) return false;
  // threshold < 1 keeps the moved clusters (<= floor(average * threshold) rows) apart from the source clusters
  // (>= average rows), so the moves can run in parallel
  if (!(threshold < 1.0f)) fail(kBadState, "balancing threshold must be below 1");
  const uint64_t average = n / k;
  uint64_t ofst;
  do {
    f.prime_j = (f.prime_j + 1) % kStridePrimes;
    ofst = stride_prime(f.prime_j);
  } while (n % ofst == 0);
  const uint64_t small = (uint64_t)((float)average * threshold);
  f.cnt.resize(k);
  HDG_HIP(hipMemcpyAsync(f.cnt.data(), f.d_cnt.data(), k * sizeof(unsigned long long), hipMemcpyDeviceToHost, f.s));
  HDG_HIP(hipStreamSynchronize(f.s));
  bool any = false;
  for (uint32_t l = 0; l < k && !any; ++l) any = f.cnt[l] <= small;
  if (!any) return false;
  f.lab.resize(n);
  HDG_HIP(hipMemcpyAsync(f.lab.data(), label, n * sizeof(uint32_t), hipMemcpyDeviceToHost, f.s));
  HDG_HIP(hipStreamSynchronize(f.s));
  std::vector<uint32_t> moves;
  std::vector<float> wcs;
  for (uint32_t l = 0; l < k; ++l) {
    if (f.cnt[l] > small) continue;
    do {
      f.cursor = (f.cursor + ofst) % n;
    } while (f.cnt[f.lab[f.cursor]] < average);
    moves.insert(moves.end(), {l, f.lab[f.cursor], (uint32_t)f.cursor});
    wcs.push_back(std::min((float)f.cnt[l], kAdjustWeight));
  }
  upload(f.d_moves, moves, f.s);
  upload(f.d_wc, wcs, f.s);
  hipLaunchKernelGGL(adjust_move_kernel, dim3((unsigned)ceil_div(w, 64), (unsigned)wcs.size()), dim3(64), 0, f.s,
                     centers, w, f.d_moves.as<uint32_t>(), f.d_wc.as<float>(), X, ld);
  HDG_HIP(hipGetLastError());
  return true;
}

void bal_em(BalancedFit& f, const float* X, uint64_t n, uint64_t ld, uint32_t w, uint32_t k, uint32_t n_iters,
            uint32_t pullback, float threshold, double scale, float* centers, uint32_t* label) {
  uint32_t counter = pullback;
  for (uint32_t iter = 0; iter < n_iters; ++iter) {
    if (iter > 0 && bal_adjust(f, X, n, ld, w, k, label, threshold, centers)) {
      if (counter++ >= pullback) {
        counter -= pullback;
        ++n_iters;
      }
    }
    bal_predict(f, X, n, ld, w, k, centers, label);
    bal_centers(f, X, n, ld, w, k, label, scale, centers);
  }
}

// Train k clusters on X from labels i mod k (pullback 2, threshold 0.25).
void bal_build_clusters(BalancedFit& f, const float* X, uint64_t n, uint64_t ld, uint32_t w, uint32_t k,
                        uint32_t n_iters, float* centers, uint32_t* label) {
  const double scale = fixed_point_scale(X, n, ld, w, f.s);
  if (n > 0) {
    hipLaunchKernelGGL(mod_label_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, f.s, label, n, k);
    HDG_HIP(hipGetLastError());
  }
  bal_centers(f, X, n, ld, w, k, label, scale, centers);
  bal_em(f, X, n, ld, w, k, n_iters, 2, 0.25f, scale, centers, label);
}

}  // namespace

// Fine cluster counts per mesocluster: proportional to the mesocluster sizes, rounded half up, at least one
// for a non-empty mesocluster while leaving one for each non-empty one after it; the last takes the rest.
std::vector<uint32_t> arrange_fine_clusters(uint32_t k, const std::vector<uint64_t>& meso_sizes, uint64_t n) {
  const size_t nm = meso_sizes.size();
  std::vector<uint32_t> nums(nm, 0);
  uint64_t lists_rem = k, rows_rem = n, nonempty_rem = 0;
  for (uint64_t sz : meso_sizes) nonempty_rem += sz > 0;
  for (size_t i = 0; i < nm; ++i) {
    if (i + 1 < nm) {
      if (meso_sizes[i] > 0) {
        --nonempty_rem;
        uint64_t sh = (uint64_t)((double)(lists_rem * meso_sizes[i]) / (double)rows_rem + 0.5);
        sh = std::min<uint64_t>(sh, lists_rem - nonempty_rem);
        nums[i] = (uint32_t)std::max<uint64_t>(sh, 1);
      }
    } else {
      nums[i] = (uint32_t)lists_rem;
    }
    lists_rem -= nums[i];
    rows_rem -= meso_sizes[i];
  }
  return nums;
}

void kmeans_balanced(const float* X, uint64_t n, uint32_t w, uint32_t k, uint32_t n_iters,
                     std::vector<float>& centers, std::vector<uint32_t>& labels, hipStream_t s) {
  if (k == 0 || n < k) fail(kInvalidArgument, "balanced k-means needs at least as many rows as clusters");
  BalancedFit f;
  f.s = s;
  // mesoclusters over all rows
  const uint32_t nm = std::min<uint32_t>(k, (uint32_t)(std::sqrt((double)k) + 0.5));
  DevBuf d_meso_c, d_meso_l;
  d_meso_c.reserve((size_t)nm * w * sizeof(float));
  d_meso_l.reserve(n * sizeof(uint32_t));
  bal_build_clusters(f, X, n, w, w, nm, n_iters, d_meso_c.as<float>(), d_meso_l.as<uint32_t>());
  std::vector<uint32_t> meso_l(n);
  HDG_HIP(hipMemcpyAsync(meso_l.data(), d_meso_l.data(), n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HDG_HIP(hipStreamSynchronize(s));
  std::vector<uint64_t> meso_sizes(nm, 0);
  for (uint32_t v : meso_l) ++meso_sizes[v];
  const std::vector<uint32_t> nums = arrange_fine_clusters(k, meso_sizes, n);
  // fine clusters inside each mesocluster (rows in ascending order)
  DevBuf d_c, d_l, d_sub, d_sub_l;
  d_c.reserve((size_t)k * w * sizeof(float));
  HDG_HIP(hipMemsetAsync(d_c.data(), 0, (size_t)k * w * sizeof(float), s));
  d_l.reserve(n * sizeof(uint32_t));
  uint64_t done = 0;
  for (uint32_t i = 0; i < nm; ++i) {
    std::vector<uint32_t> rows;
    for (uint64_t j = 0; j < n; ++j)
      if (meso_l[j] == i) rows.push_back((uint32_t)j);
    if (rows.empty() || nums[i] == 0) {  // an empty last mesocluster leaves its share at zero centers
      done += nums[i];
      continue;
    }
    gather_rows_dev(X, w, rows, w, d_sub, s);
    d_sub_l.reserve(rows.size() * sizeof(uint32_t));
    bal_build_clusters(f, d_sub.as<float>(), rows.size(), w, w, nums[i], n_iters, d_c.as<float>() + done * w,
                       d_sub_l.as<uint32_t>());
    done += nums[i];
  }
  // balancing EM over all rows with all k centers, then the final assignment
  const double scale = fixed_point_scale(X, n, w, w, s);
  bal_em(f, X, n, w, w, k, std::max<uint32_t>(n_iters / 10, 2), 5, 0.2f, scale, d_c.as<float>(),
         d_l.as<uint32_t>());
  bal_predict(f, X, n, w, w, k, d_c.as<float>(), d_l.as<uint32_t>());
  centers.resize((size_t)k * w);
  labels.resize(n);
  HDG_HIP(hipMemcpyAsync(centers.data(), d_c.data(), (size_t)k * w * sizeof(float), hipMemcpyDeviceToHost, s));
  HDG_HIP(hipMemcpyAsync(labels.data(), d_l.data(), n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HDG_HIP(hipStreamSynchronize(s));
}

namespace {

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

void build_bucket_grid(IndexState& st, hipStream_t s) {
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
  // coarse occupancy grid: 2^(12 / n_valid) coarse cells per dim (at most the dim's cells)
  const int gb = std::max(1, 12 / nv);
  for (int j = 0; j < nv; ++j) st.sg[j] = std::min<uint32_t>(1u << gb, st.Mv[j]);
  st.sstride[nv - 1] = 1;
  for (int j = nv - 2; j >= 0; --j) st.sstride[j] = st.sstride[j + 1] * (st.sg[j + 1] + 1);
  BucketArgs a{};
  a.nv = nv;
  for (int j = 0; j < nv; ++j) {
    a.Mv[j] = st.Mv[j];
    a.nb[j] = st.nb[j];
    a.sg[j] = st.sg[j];
    a.bstride[j] = st.bstride[j];
    a.sstride[j] = st.sstride[j];
  }
  const unsigned blocks = (unsigned)ceil_div(N, 256);
  // entries: z-pool positions stably sorted by bucket (ascending position inside a bucket)
  {
    DevBuf bid, pos, bid_sorted;
    bid.reserve(N * 4);
    pos.reserve(N * 4);
    bid_sorted.reserve(N * 4);
    st.epos.reserve(N * 4);
    hipLaunchKernelGGL(bucket_id_kernel, dim3(blocks), dim3(256), 0, s, st.cells.as<uint32_t>(), N, a,
                       bid.as<uint32_t>(), pos.as<uint32_t>());
    HDG_HIP(hipGetLastError());
    sort_pairs(bid.as<uint32_t>(), bid_sorted.as<uint32_t>(), pos.as<uint32_t>(), st.epos.as<uint32_t>(), N,
               key_bits(st.Nb), s);
    st.bstart.reserve((st.Nb + 1) * 4);
    hipLaunchKernelGGL(lower_bounds_kernel, dim3((unsigned)ceil_div(st.Nb + 1, 256)), dim3(256), 0, s,
                       bid_sorted.as<uint32_t>(), N, st.Nb, st.bstart.as<uint32_t>());
    HDG_HIP(hipGetLastError());
    // compact entries: u16 cells when every valid dim has at most 65536 cells; sub-cluster and position in one
    // word when their bits fit (the walk then reads 4 bytes fewer per tested entry and lists 4 bytes per point)
    st.cells16 = true;
    for (int j = 0; j < nv; ++j) st.cells16 = st.cells16 && st.Mv[j] <= 65536u;
    const int sub_bits = key_bits(st.CS), off_bits = key_bits(st.max_seg_len);
    st.entries_packed = sub_bits + off_bits <= 32 && off_bits < 32;
    st.pk_shift = st.entries_packed ? (uint32_t)off_bits : 0u;
    const int shift = st.entries_packed ? off_bits : -1;
    st.ecell.reserve((uint64_t)nv * N * (st.cells16 ? 2 : 4));
    st.esub.reserve(N * 4);
    if (st.cells16)
      hipLaunchKernelGGL(entry_fill_kernel<uint16_t>, dim3(blocks), dim3(256), 0, s, st.cells.as<uint32_t>(),
                         st.seg_off.as<uint32_t>(), st.CS, N, nv, st.epos.as<uint32_t>(), shift,
                         st.ecell.as<uint16_t>(), st.esub.as<uint32_t>());
    else
      hipLaunchKernelGGL(entry_fill_kernel<uint32_t>, dim3(blocks), dim3(256), 0, s, st.cells.as<uint32_t>(),
                         st.seg_off.as<uint32_t>(), st.CS, N, nv, st.epos.as<uint32_t>(), shift,
                         st.ecell.as<uint32_t>(), st.esub.as<uint32_t>());
    HDG_HIP(hipGetLastError());
    HDG_HIP(hipStreamSynchronize(s));
    if (st.entries_packed) st.epos.release();
  }
  // coarse occupancy sums per sub-cluster, prefix-summed along each dim
  const uint64_t per = st.sstride[0] * (st.sg[0] + 1);
  if (per * st.CS > (1ull << 26)) return;  // more than 256 MiB: the dense count scans without it
  st.sat.reserve(per * st.CS * 4);
  HDG_HIP(hipMemsetAsync(st.sat.data(), 0, per * st.CS * 4, s));
  hipLaunchKernelGGL(sat_count_kernel, dim3(blocks), dim3(256), 0, s, st.cells.as<uint32_t>(),
                     st.seg_off.as<uint32_t>(), st.CS, N, a, per, st.sat.as<uint32_t>());
  HDG_HIP(hipGetLastError());
  for (int j = 0; j < nv; ++j) {
    const uint64_t wj = st.sg[j] + 1, lines = st.CS * (per / wj);
    hipLaunchKernelGGL(sat_prefix_kernel, dim3((unsigned)ceil_div(lines, 256)), dim3(256), 0, s,
                       st.sat.as<uint32_t>(), lines, per, wj, st.sstride[j]);
    HDG_HIP(hipGetLastError());
  }
  HDG_HIP(hipStreamSynchronize(s));
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

  // per-stage wall times at log level 1 (HDG_LOG=1)
  auto t_last = std::chrono::steady_clock::now();
  auto stage_done = [&](const char* what) {
    if (log_level() < 1) return;
    HDG_HIP(hipStreamSynchronize(s));
    const auto now = std::chrono::steady_clock::now();
    log_line(1, "build %s: %.1f ms", what, std::chrono::duration<double, std::milli>(now - t_last).count());
    t_last = now;
  };

  // ---- B1: grid shape from the labels (host copy) ----
  std::vector<float> h_lab(n * l);
  HDG_HIP(hipMemcpyAsync(h_lab.data(), labels, n * l * sizeof(float), hipMemcpyDeviceToHost, s));
  HDG_HIP(hipStreamSynchronize(s));
  st_.grid = compute_grid_shape(h_lab.data(), n, l, cfg_.max_grids, cfg_.label_mask);
  const GridShape& g = st_.grid;
  stage_done("grid shape");

  // ---- B6a: PQ codebooks per subspace (raw vectors, product-quantization.cu:182-211) ----
  std::vector<float> pq_centers(M * K * pq_len, 0.0f);
  DevBuf d_codes_orig;  // u8 [n][M] original order
  d_codes_orig.reserve(n * M);
  {
    std::vector<uint32_t> rows = sample_rows(n, std::max<uint64_t>(cfg_.pq_train, K), cfg_.seed ^ 0x5151);
    DevBuf d_train, d_sub;
    gather_rows_dev(dataset, d, rows, (uint32_t)d, d_train, s);
    d_sub.reserve(rows.size() * pq_len * sizeof(float));
    for (uint64_t m = 0; m < M; ++m) {
      const uint64_t off = m * pq_len;
      const uint32_t w = (uint32_t)std::min<uint64_t>(pq_len, d - std::min(d, off));
      if (w == 0) continue;  // an empty subspace keeps zero centers
      hipLaunchKernelGGL(slice_kernel, dim3((unsigned)ceil_div(rows.size() * w, 256)), dim3(256), 0, s,
                         d_train.as<float>(), (uint64_t)rows.size(), d, off, w, d_sub.as<float>());
      HDG_HIP(hipGetLastError());
      std::vector<float> cw;
      kmeans(d_sub.as<float>(), rows.size(), w, w, (uint32_t)K, (int)cfg_.kmeans_iters,
             cfg_.seed * 1315423911ull + m, cw, s);
      for (uint64_t c = 0; c < K; ++c)
        for (uint32_t t = 0; t < w; ++t) pq_centers[(m * K + c) * pq_len + t] = cw[c * w + t];
    }
    // encode all n rows: the one-pass encoder when a subspace fits its registers, else per subspace
    std::vector<float> pq_norm(M * K);
    for (uint64_t e = 0; e < M * K; ++e) {
      double a = 0;
      for (uint64_t t = 0; t < pq_len; ++t) a += (double)pq_centers[e * pq_len + t] * pq_centers[e * pq_len + t];
      pq_norm[e] = (float)a;
    }
    DevBuf d_pqc, d_pqn;
    upload(d_pqc, pq_centers, s);
    upload(d_pqn, pq_norm, s);
    const size_t lds = (K * pq_len + K) * sizeof(float);
    const dim3 grid((unsigned)ceil_div(n, 256));
    auto enc = [&](auto kern) {
      hipLaunchKernelGGL(kern, grid, dim3(256), lds, s, dataset, n, d, (uint32_t)M, (uint32_t)K, (uint32_t)pq_len,
                         d_pqc.as<float>(), d_pqn.as<float>(), d_codes_orig.as<uint8_t>());
      HDG_HIP(hipGetLastError());
    };
    if (pq_len <= 4) enc(pq_encode_kernel<4>);
    else if (pq_len <= 8) enc(pq_encode_kernel<8>);
    else if (pq_len <= 16) enc(pq_encode_kernel<16>);
    else {
      DevBuf d_all, d_lab32;
      d_all.reserve(n * pq_len * sizeof(float));
      for (uint64_t m = 0; m < M; ++m) {
        const uint64_t off = m * pq_len;
        const uint32_t w = (uint32_t)std::min<uint64_t>(pq_len, d - std::min(d, off));
        std::vector<float> cw(K * (uint64_t)std::max<uint32_t>(w, 1), 0.0f);
        if (w > 0) {
          for (uint64_t c = 0; c < K; ++c)
            for (uint32_t t = 0; t < w; ++t) cw[c * w + t] = pq_centers[(m * K + c) * pq_len + t];
          hipLaunchKernelGGL(slice_kernel, dim3((unsigned)ceil_div(n * w, 256)), dim3(256), 0, s, dataset, n, d, off,
                             w, d_all.as<float>());
          HDG_HIP(hipGetLastError());
          assign_all(d_all.as<float>(), n, w, w, cw, (uint32_t)K, d_lab32, s);
        } else {
          d_lab32.reserve(n * sizeof(uint32_t));
          HDG_HIP(hipMemsetAsync(d_lab32.data(), 0, n * sizeof(uint32_t), s));
        }
        hipLaunchKernelGGL(code_column_kernel, grid, dim3(256), 0, s, d_lab32.as<uint32_t>(), n, (uint32_t)M,
                           (uint32_t)m, d_codes_orig.as<uint8_t>());
        HDG_HIP(hipGetLastError());
      }
    }
    HDG_HIP(hipStreamSynchronize(s));
  }
  stage_done("PQ codebooks and codes");

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
  stage_done("fine k-means and assignment");

  // ---- B6c: group sub-centers into C clusters with the balanced k-means (kmeans_balanced, :466-476) ----
  std::vector<float> coarse;
  std::vector<uint32_t> group;
  {
    DevBuf d_fine;
    upload(d_fine, fine, s);
    kmeans_balanced(d_fine.as<float>(), CS, (uint32_t)d, (uint32_t)C, (uint32_t)cfg_.balance_iters, coarse, group, s);
  }
  stage_done("sub-center grouping");

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

  // ---- B2/B5 on the device: cells and Morton codes; z-pool = per sub-cluster ascending (zcode, id) ----
  // Two stable radix sorts: (zcode -> id) with ids ascending, then (sub-cluster -> id) on that order.
  const int nv = g.n_valid;
  st_.h_seg_off.assign(CS + 1, 0);
  {
    GridArgs ga{};
    ga.l = l;
    ga.max_bits = g.max_bits;
    ga.valid_map = g.valid_map;
    ga.n_valid = nv;
    for (int i = 0; i < kMaxDims; ++i) {
      ga.dims[i] = g.dims[i];
      ga.shuffle[i] = g.shuffle[i];
      ga.valid_dims[i] = g.valid_dims[i];
      ga.left[i] = g.left[i];
      ga.span[i] = g.span[i];
    }
    DevBuf d_pos_of_fine, cells_orig, zc, zc_s, seg, seg_s, ids, ids_s;
    upload(d_pos_of_fine, pos_of_fine, s);
    cells_orig.reserve((uint64_t)nv * n * 4);
    zc.reserve(n * 8);
    zc_s.reserve(n * 8);
    seg.reserve(n * 4);
    seg_s.reserve(n * 4);
    ids.reserve(n * 4);
    ids_s.reserve(n * 4);
    const dim3 grid((unsigned)ceil_div(n, 256));
    hipLaunchKernelGGL(cells_morton_kernel, grid, dim3(256), 0, s, labels, n, ga, d_fine_label.as<uint32_t>(),
                       d_pos_of_fine.as<uint32_t>(), cells_orig.as<uint32_t>(), zc.as<uint64_t>(), seg.as<uint32_t>(),
                       ids.as<uint32_t>());
    HDG_HIP(hipGetLastError());
    const int zbits = std::min(64, nv * g.max_bits);
    sort_pairs(zc.as<uint64_t>(), zc_s.as<uint64_t>(), ids.as<uint32_t>(), ids_s.as<uint32_t>(), n, zbits, s);
    // seg in (zcode, id) order, then the stable sort by sub-cluster: ids_s -> st_.zid
    hipLaunchKernelGGL(gather_kernel<uint32_t>, grid, dim3(256), 0, s, seg.as<uint32_t>(), ids_s.as<uint32_t>(), n,
                       ids.as<uint32_t>());  // ids now holds the gathered sub-cluster keys
    HDG_HIP(hipGetLastError());
    st_.zid.reserve(n * 4);
    sort_pairs(ids.as<uint32_t>(), seg_s.as<uint32_t>(), ids_s.as<uint32_t>(), st_.zid.as<uint32_t>(), n,
               key_bits(CS), s);
    st_.seg_off.reserve((CS + 1) * 4);
    hipLaunchKernelGGL(lower_bounds_kernel, dim3((unsigned)ceil_div(CS + 1, 256)), dim3(256), 0, s,
                       seg_s.as<uint32_t>(), n, CS, st_.seg_off.as<uint32_t>());
    HDG_HIP(hipGetLastError());
    // per-point arrays in z-pool order
    st_.zcode.reserve(n * 8);
    hipLaunchKernelGGL(gather_kernel<uint64_t>, grid, dim3(256), 0, s, zc.as<uint64_t>(), st_.zid.as<uint32_t>(), n,
                       st_.zcode.as<uint64_t>());
    HDG_HIP(hipGetLastError());
    st_.cells.reserve((uint64_t)nv * n * 4);
    for (int j = 0; j < nv; ++j) {
      hipLaunchKernelGGL(gather_kernel<uint32_t>, grid, dim3(256), 0, s, cells_orig.as<uint32_t>() + (uint64_t)j * n,
                         st_.zid.as<uint32_t>(), n, st_.cells.as<uint32_t>() + (uint64_t)j * n);
      HDG_HIP(hipGetLastError());
    }
    st_.lab.reserve(n * l * sizeof(float));
    hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)ceil_div(n, 4)), dim3(64, 4), 0, s,
                       reinterpret_cast<const uint32_t*>(labels), (uint64_t)l, st_.zid.as<uint32_t>(), n, (uint32_t)l,
                       st_.lab.as<uint32_t>());
    HDG_HIP(hipGetLastError());
    st_.vec.reserve(n * d * sizeof(float));
    hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)ceil_div(n, 4)), dim3(64, 4), 0, s,
                       reinterpret_cast<const uint32_t*>(dataset), d, st_.zid.as<uint32_t>(), n, (uint32_t)d,
                       st_.vec.as<uint32_t>());
    HDG_HIP(hipGetLastError());
    st_.codes.reserve(n * M);
    hipLaunchKernelGGL(gather_bytes_kernel, dim3((unsigned)ceil_div(n * M, 256)), dim3(256), 0, s,
                       d_codes_orig.as<uint8_t>(), st_.zid.as<uint32_t>(), n, (uint32_t)M, st_.codes.as<uint8_t>());
    HDG_HIP(hipGetLastError());
    // host copies for save and the per-query limits
    std::vector<uint32_t> so(CS + 1);
    st_.h_zid.resize(n);
    st_.h_zcode.resize(n);
    HDG_HIP(hipMemcpyAsync(so.data(), st_.seg_off.data(), (CS + 1) * 4, hipMemcpyDeviceToHost, s));
    HDG_HIP(hipMemcpyAsync(st_.h_zid.data(), st_.zid.data(), n * 4, hipMemcpyDeviceToHost, s));
    HDG_HIP(hipMemcpyAsync(st_.h_zcode.data(), st_.zcode.data(), n * 8, hipMemcpyDeviceToHost, s));
    HDG_HIP(hipStreamSynchronize(s));
    for (uint64_t i = 0; i <= CS; ++i) st_.h_seg_off[i] = so[i];
    for (uint64_t i = 0; i < CS; ++i)
      st_.max_seg_len = std::max<uint32_t>(st_.max_seg_len, (uint32_t)(so[i + 1] - so[i]));
  }
  stage_done("cells, Morton codes, z-pool sort and gathers");
  build_bucket_grid(st_, s);
  stage_done("bucket grid and occupancy sums");
  prepare_label_words(st_, h_lab, s);
  upload(st_.pq_centers, pq_centers, s);
  upload(st_.centers, coarse, s);
  prepare_mixed_centers(st_, coarse, sub_centers, s);
  upload(st_.sub_centers, sub_centers, s);
  upload(st_.sub_off, st_.h_sub_off, s);
  upload(st_.cluster_of_sub, cluster_of_sub, s);
  prepare_blocked_codes(st_, s);
  HDG_HIP(hipStreamSynchronize(s));
  stage_done("centers and blocked codes");
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

namespace {
// One thread per point: its label dims as offsets from the column minima, packed into one word.
__global__ void label_words_kernel(const float* __restrict__ lab, uint64_t n, int l, LabelWordLayout lay,
                                   uint64_t* __restrict__ out) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint64_t w = 0;
  for (int j = 0; j < l; ++j) w |= ((uint64_t)lab[p * l + j] - lay.minv[j]) << lay.off[j];
  out[p] = w;
}
}  // namespace

void prepare_label_words(IndexState& st, const std::vector<float>& h_lab, hipStream_t s) {
  const int l = st.l;
  st.lab_words = false;
  st.lw = LabelWordLayout{};
  if (l < 1 || l > kMaxDims) {  // the grid shape refuses such l before this; no words then
    st.labels_integral = std::all_of(h_lab.begin(), h_lab.end(), [](float v) { return v == std::floor(v); });
    return;
  }
  float mn[kMaxDims], mx[kMaxDims];
  for (int j = 0; j < l; ++j) {
    mn[j] = FLT_MAX;
    mx[j] = -FLT_MAX;
  }
  bool integral = true;
  const uint64_t n = h_lab.size() / (l > 0 ? l : 1);
  for (uint64_t p = 0; p < n; ++p) {
    const float* r = h_lab.data() + p * l;
    for (int j = 0; j < l; ++j) {
      const float v = r[j];
      integral = integral && v == std::floor(v);
      mn[j] = v < mn[j] ? v : mn[j];
      mx[j] = v > mx[j] ? v : mx[j];
    }
  }
  st.labels_integral = integral;
  if (!integral || n == 0) return;
  // the layout is built aside and taken only when every column fits
  LabelWordLayout lay{};
  uint32_t off = 0;
  for (int j = 0; j < l; ++j) {
    if (!(mn[j] >= 0.0f) || !(mx[j] <= 9007199254740992.0f)) return;  // 2^53: exact in double
    const uint64_t lo = (uint64_t)mn[j], range = (uint64_t)mx[j] - lo;
    if (range >= (1ull << 31)) return;
    uint32_t bits = 1;
    while (bits < 31 && (range >> bits) != 0) ++bits;
    if (off + bits + 1 > 64) return;
    lay.off[j] = off;
    lay.mask[j] = (uint32_t)((1ull << bits) - 1);
    lay.minv[j] = lo;
    lay.guard |= 1ull << (off + bits);
    off += bits + 1;
  }
  st.lw = lay;
  st.labw.reserve(st.N * sizeof(uint64_t));
  hipLaunchKernelGGL(label_words_kernel, dim3((unsigned)((st.N + 255) / 256)), dim3(256), 0, s, st.lab.as<float>(),
                     st.N, l, st.lw, st.labw.as<uint64_t>());
  HDG_HIP(hipGetLastError());
  st.lw_dev.reserve(sizeof(LabelWordLayout));
  HDG_HIP(hipMemcpyAsync(st.lw_dev.data(), &st.lw, sizeof(LabelWordLayout), hipMemcpyHostToDevice, s));
  HDG_HIP(hipStreamSynchronize(s));  // the host copy of the layout may change before the stream runs
  st.lab_words = true;
}

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
