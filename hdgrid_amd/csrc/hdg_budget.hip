// Learned per-query probe budget for PRECISION 1 (BUDGET_MODEL = 1).
//
// Budget ladder: level i keeps S[i] sub-clusters (growing geometrically) of all non-zero clusters.  A training
// query's label is the smallest level at which the labelling path (BUDGET_LABELS 0: the plain reference-exact
// path; 1: the mixed path) reaches the target recall@k against the ground truth supplied by the caller.
// Model: 48 features -> 64 ReLU -> 64 ReLU -> 1 (predicted level).  Features: log1p of the filter width of
// each label dim (8), log1p of the grid-cell extent of each valid dim (8), the 32 smallest BF16 coarse-centroid
// scores among the query's non-zero clusters, plus |q|^2 (32); standardised with training mean / std.
// Training: BF16 operands (inputs, weights and activations rounded to bf16), FP32 accumulation, FP32
// master weights, Adam, MSE loss.  Inference adds a margin: the smallest one at which the deployed path (mixed
// precision with the model) reaches the target mean recall plus one standard error of that mean on a held-out
// share of the training queries; a non-finite prediction falls back to the fixed budget (the highest level).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <numeric>
#include <random>

#include "hdg_budget.hpp"
#include "hdg_fast.hpp"
#include "hdg_index.hpp"

namespace hdg {

namespace {

constexpr int kF = kBudgetFeatures;  // 48
constexpr int kH = 64;
constexpr int kP1 = kF * kH, kP2 = kH * kH;
constexpr int kOffW1 = 0, kOffB1 = kP1, kOffW2 = kP1 + kH, kOffB2 = kP1 + kH + kP2, kOffW3 = kOffB2 + kH,
              kOffB3 = kOffW3 + kH;
constexpr int kParams = kOffB3 + 1;  // 7361

__device__ __forceinline__ float bf16r(float v) {
  uint32_t u = __float_as_uint(v);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return __uint_as_float(u & 0xFFFF0000u);
}

// Features of one query (unstandardised).
__device__ void query_features(uint64_t qi, int l, const float* ranges, const uint32_t* box_w, int n_valid,
                               uint64_t qn, const float* feat_dist, float* f) {
  for (int i = 0; i < 8; ++i) {
    float v = 0.0f;
    if (i < l) {
      float w = ranges[qi * 2 * l + 2 * i + 1] - ranges[qi * 2 * l + 2 * i];
      w = w > 0.0f ? (w < 1e12f ? w : 1e12f) : 0.0f;
      v = log1pf(w);
    }
    f[i] = v;
  }
  for (int j = 0; j < 8; ++j) f[8 + j] = j < n_valid ? log1pf((float)box_w[(uint64_t)j * qn + qi] + 1.0f) : 0.0f;
  for (int j = 0; j < 32; ++j) f[16 + j] = feat_dist[qi * kFeatDists + j];
}

__global__ void features_kernel(uint64_t qn, int l, const float* __restrict__ ranges,
                                const uint32_t* __restrict__ box_w, int n_valid, const float* __restrict__ feat_dist,
                                float* __restrict__ out) {
  const uint64_t qi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (qi >= qn) return;
  float f[kF];
  query_features(qi, l, ranges, box_w, n_valid, qn, feat_dist, f);
  for (int i = 0; i < kF; ++i) out[qi * kF + i] = f[i];
}

// Forward of one sample with bf16-rounded operands; stores pre-activations when z1/z2 are given.
__device__ float mlp_forward(const float* w, const float* x, float* z1, float* z2) {
  float h1[kH], h2[kH];
  for (int j = 0; j < kH; ++j) {
    float a = w[kOffB1 + j];
    for (int i = 0; i < kF; ++i) a += bf16r(x[i]) * w[kOffW1 + i * kH + j];
    if (z1) z1[j] = a;
    h1[j] = bf16r(a > 0.0f ? a : 0.0f);
  }
  for (int j = 0; j < kH; ++j) {
    float a = w[kOffB2 + j];
    for (int i = 0; i < kH; ++i) a += h1[i] * w[kOffW2 + i * kH + j];
    if (z2) z2[j] = a;
    h2[j] = bf16r(a > 0.0f ? a : 0.0f);
  }
  float y = w[kOffB3];
  for (int i = 0; i < kH; ++i) y += h2[i] * w[kOffW3 + i];
  return y;
}

// Inference: features -> standardise -> MLP -> level -> sub-cluster budget; trims the query's chosen
// sub-clusters (sorted by score) to the budget.
// 16 lanes per query: each lane owns 4 hidden units of each layer (every unit's sum runs over its inputs
// in order, as in mlp_forward), the output sum is taken by the group's first lane.  A workgroup of 256
// threads serves 16 queries and keeps the bf16 weights in LDS.
constexpr int kInferQ = 16;
constexpr uint64_t kInferWgs = 1024;  // workgroups of the inference (four fit a CU; each stages the weights once)

__global__ __launch_bounds__(256) void infer_kernel(uint64_t qn, int l, const float* __restrict__ ranges,
                                                    const uint32_t* __restrict__ box_w, int n_valid,
                                                    const float* __restrict__ feat_dist,
                                                    const float* __restrict__ wq, const float* __restrict__ mean,
                                                    const float* __restrict__ inv_std, float margin, int levels,
                                                    const uint32_t* __restrict__ ladT,
                                                    const uint32_t* __restrict__ ladS,
                                                    uint32_t* __restrict__ n_chosen_s, uint32_t* __restrict__ bT,
                                                    uint32_t* __restrict__ bS,
                                                    const uint32_t* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  __shared__ __align__(16) float w[kParams];
  __shared__ float xs[kInferQ][kF];
  __shared__ float h1s[kInferQ][kH];
  __shared__ float h2s[kInferQ][kH];
  {
    // the weights as 16-byte pieces, several in flight per thread
    constexpr int n4 = kParams / 4;
    const float4* w4 = reinterpret_cast<const float4*>(wq);
#pragma unroll 8
    for (int i = threadIdx.x; i < n4; i += blockDim.x) reinterpret_cast<float4*>(w)[i] = w4[i];
    for (int i = n4 * 4 + threadIdx.x; i < kParams; i += blockDim.x) w[i] = wq[i];
  }
  const int g = threadIdx.x & 15, qg = threadIdx.x >> 4;
  // persistent over groups of kInferQ queries: the weights are staged once per workgroup
  const uint64_t groups = (qn + kInferQ - 1) / kInferQ;
  for (uint64_t grp = blockIdx.x; grp < groups; grp += gridDim.x) {
    const uint64_t qi = grp * kInferQ + qg;
    const bool live = qi < qn;
    __syncthreads();  // the previous group is done with xs / h1s / h2s (and the weights are staged)
    if (live) {
      // the query's 48 features (as query_features), three per lane
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int i = g + 16 * k;
        float v = 0.0f;
        if (i < 8) {
          if (i < l) {
            float wd = ranges[qi * 2 * l + 2 * i + 1] - ranges[qi * 2 * l + 2 * i];
            wd = wd > 0.0f ? (wd < 1e12f ? wd : 1e12f) : 0.0f;
            v = log1pf(wd);
          }
        } else if (i < 16) {
          const int j = i - 8;
          v = j < n_valid ? log1pf((float)box_w[(uint64_t)j * qn + qi] + 1.0f) : 0.0f;
        } else {
          v = feat_dist[qi * kFeatDists + (i - 16)];
        }
        xs[qg][i] = bf16r((v - mean[i]) * inv_std[i]);
      }
    }
    __syncthreads();
    // hidden units 4g .. 4g+3 of each layer: four independent chains, each summed in input order, their
    // weights one 16-byte LDS read per input
    if (live) {
      float a[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) a[u] = w[kOffB1 + g * 4 + u];
#pragma unroll 8
      for (int i = 0; i < kF; ++i) {
        const float x = xs[qg][i];
        const float4 wv = *reinterpret_cast<const float4*>(&w[kOffW1 + i * kH + g * 4]);
        a[0] += x * wv.x;
        a[1] += x * wv.y;
        a[2] += x * wv.z;
        a[3] += x * wv.w;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) h1s[qg][g * 4 + u] = bf16r(a[u] > 0.0f ? a[u] : 0.0f);
    }
    __syncthreads();
    if (live) {
      float a[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) a[u] = w[kOffB2 + g * 4 + u];
#pragma unroll 8
      for (int i = 0; i < kH; ++i) {
        const float x = h1s[qg][i];
        const float4 wv = *reinterpret_cast<const float4*>(&w[kOffW2 + i * kH + g * 4]);
        a[0] += x * wv.x;
        a[1] += x * wv.y;
        a[2] += x * wv.z;
        a[3] += x * wv.w;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) h2s[qg][g * 4 + u] = bf16r(a[u] > 0.0f ? a[u] : 0.0f);
    }
    __syncthreads();
    if (live && g == 0) {
      float y = w[kOffB3];
#pragma unroll 16
      for (int i = 0; i < kH; ++i) y += h2s[qg][i] * w[kOffW3 + i];
      int lev = levels - 1;  // fallback: the fixed budget
      if (isfinite(y)) {
        float t = ceilf(y + margin);
        lev = t < 0.0f ? 0 : (t > (float)(levels - 1) ? levels - 1 : (int)t);
      }
      bT[qi] = ladT[lev];
      const uint32_t S = ladS[lev];
      bS[qi] = S;
      if (n_chosen_s[qi] > S) n_chosen_s[qi] = S;
    }
  }
}


// Training forward + backward over a mini-batch, kTrainTile samples per workgroup of 256 threads.  Phase 1:
// four threads per sample, thread t owning hidden units 16t .. 16t + 15 of each layer (every unit's sum runs
// over its inputs in order, as in mlp_forward), leave the sample's bf16 inputs, activations and back-propagated
// errors in LDS; the output and its error are one thread's sequential sum.  Phase 2: each thread takes
// parameters and sums their gradient over the tile's samples in order into the workgroup's row of gpart
// [workgroups][kParams]; the tile's squared errors / bsz are summed in sample order into lpart[workgroup].
// grad_reduce_kernel then adds the rows in workgroup order, so a step's gradient (and the whole fit) does not
// depend on which workgroup finishes first.  Samples past the batch carry zero errors.  LDS: the weights plus
// seven tile arrays (dynamic).
constexpr int kTrainTile = 64;
constexpr int kTrainThreads = 256;
constexpr int kUnitsPer = kH / (kTrainThreads / kTrainTile);  // 16
constexpr int kParamsPad = (kParams + 3) / 4 * 4;
constexpr size_t kTrainLds = (size_t)(kParamsPad + kTrainTile * (kF + 4 * kH) + 2 * kTrainTile) * sizeof(float);

__global__ __launch_bounds__(256) void train_step_kernel(const float* __restrict__ X, const float* __restrict__ Y,
                                                         const uint32_t* __restrict__ idx, uint32_t bsz,
                                                         const float* __restrict__ wq, float* __restrict__ gpart,
                                                         float* __restrict__ lpart) {
  extern __shared__ __align__(16) float tsm[];
  float* w = tsm;                                 // [kParams]
  float* xs = w + kParamsPad;                     // [tile][kF] bf16-rounded inputs
  float* h1s = xs + kTrainTile * kF;              // [tile][kH]
  float* h2s = h1s + kTrainTile * kH;             // [tile][kH]
  float* g2s = h2s + kTrainTile * kH;             // [tile][kH] errors at layer-2 pre-activations
  float* g1s = g2s + kTrainTile * kH;             // [tile][kH] errors at layer-1 pre-activations
  float* gys = g1s + kTrainTile * kH;             // [tile] output errors
  float* les = gys + kTrainTile;                  // [tile] squared errors / bsz
  const int tid = threadIdx.x;
  for (int i = tid; i < kParams; i += kTrainThreads) w[i] = wq[i];
  const int sl = tid >> 2, part = tid & 3;        // sample of the tile, quarter of the units
  const uint32_t b = blockIdx.x * kTrainTile + (uint32_t)sl;
  const bool valid = b < bsz;
  const uint32_t smp = valid ? idx[b] : 0u;
  for (int f = part; f < kF; f += 4) xs[sl * kF + f] = valid ? bf16r(X[(uint64_t)smp * kF + f]) : 0.0f;
  __syncthreads();
  float z1[kUnitsPer], z2[kUnitsPer];
#pragma unroll
  for (int u = 0; u < kUnitsPer; ++u) {
    const int j = part * kUnitsPer + u;
    float a = w[kOffB1 + j];
    for (int i = 0; i < kF; ++i) a += xs[sl * kF + i] * w[kOffW1 + i * kH + j];
    z1[u] = a;
    h1s[sl * kH + j] = valid ? bf16r(a > 0.0f ? a : 0.0f) : 0.0f;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kUnitsPer; ++u) {
    const int j = part * kUnitsPer + u;
    float a = w[kOffB2 + j];
    for (int i = 0; i < kH; ++i) a += h1s[sl * kH + i] * w[kOffW2 + i * kH + j];
    z2[u] = a;
    h2s[sl * kH + j] = valid ? bf16r(a > 0.0f ? a : 0.0f) : 0.0f;
  }
  __syncthreads();
  if (part == 0) {
    float gy = 0.0f, le = 0.0f;
    if (valid) {
      float y = w[kOffB3];
      for (int i = 0; i < kH; ++i) y += h2s[sl * kH + i] * w[kOffW3 + i];
      const float err = y - Y[smp];
      le = err * err / (float)bsz;
      gy = 2.0f * err / (float)bsz;
    }
    gys[sl] = gy;
    les[sl] = le;
  }
  __syncthreads();
  if (tid == 0) {
    float l = 0.0f;
    for (int t = 0; t < kTrainTile; ++t) l += les[t];
    lpart[blockIdx.x] = l;
  }
  const float gy = gys[sl];
#pragma unroll
  for (int u = 0; u < kUnitsPer; ++u) {
    const int j = part * kUnitsPer + u;
    g2s[sl * kH + j] = z2[u] > 0.0f ? bf16r(gy * w[kOffW3 + j]) : 0.0f;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kUnitsPer; ++u) {
    const int i = part * kUnitsPer + u;
    float back = 0.0f;
    for (int j = 0; j < kH; ++j) back += w[kOffW2 + i * kH + j] * g2s[sl * kH + j];
    g1s[sl * kH + i] = z1[u] > 0.0f ? bf16r(back) : 0.0f;
  }
  __syncthreads();
  // phase 2: parameter gradients summed over the tile's samples
  for (int p = tid; p < kParams; p += kTrainThreads) {
    float sum = 0.0f;
    if (p < kOffB1) {
      const int f = p / kH, i = p % kH;
      for (int t = 0; t < kTrainTile; ++t) sum += xs[t * kF + f] * g1s[t * kH + i];
    } else if (p < kOffW2) {
      const int i = p - kOffB1;
      for (int t = 0; t < kTrainTile; ++t) sum += g1s[t * kH + i];
    } else if (p < kOffB2) {
      const int i = (p - kOffW2) / kH, j = (p - kOffW2) % kH;
      for (int t = 0; t < kTrainTile; ++t) sum += h1s[t * kH + i] * g2s[t * kH + j];
    } else if (p < kOffW3) {
      const int j = p - kOffB2;
      for (int t = 0; t < kTrainTile; ++t) sum += g2s[t * kH + j];
    } else if (p < kOffB3) {
      const int j = p - kOffW3;
      for (int t = 0; t < kTrainTile; ++t) sum += gys[t] * h2s[t * kH + j];
    } else {
      for (int t = 0; t < kTrainTile; ++t) sum += gys[t];
    }
    gpart[(uint64_t)blockIdx.x * kParams + p] = sum;
  }
}

// g[p] = the workgroups' partial gradients of parameter p added in workgroup order; loss += their losses in order.
__global__ void grad_reduce_kernel(const float* __restrict__ gpart, const float* __restrict__ lpart, uint32_t nblk,
                                   float* __restrict__ g, float* __restrict__ loss) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < (uint32_t)kParams) {
    float sum = 0.0f;
    for (uint32_t b = 0; b < nblk; ++b) sum += gpart[(uint64_t)b * kParams + p];
    g[p] = sum;
  }
  if (p == 0) {
    float l = 0.0f;
    for (uint32_t b = 0; b < nblk; ++b) l += lpart[b];
    *loss += l;
  }
}

uint64_t train_step_workspace_floats(uint32_t bsz) {
  const uint64_t nblk = (bsz + kTrainTile - 1) / kTrainTile;
  return nblk * kParams + nblk;
}

// g = the batch's gradient (overwritten), *loss += the batch's mean squared error; work holds
// train_step_workspace_floats(bsz) floats.
void launch_train_step(const float* X, const float* Y, const uint32_t* idx, uint32_t bsz, const float* wq, float* g,
                       float* loss, float* work, hipStream_t s) {
  static const bool attr = [] {
    HDG_HIP(hipFuncSetAttribute((const void*)train_step_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)kTrainLds));
    return true;
  }();
  (void)attr;
  const uint32_t nblk = (bsz + kTrainTile - 1) / kTrainTile;
  float* gpart = work;
  float* lpart = work + (uint64_t)nblk * kParams;
  hipLaunchKernelGGL(train_step_kernel, dim3(nblk), dim3(kTrainThreads), kTrainLds, s, X, Y, idx, bsz, wq, gpart,
                     lpart);
  hipLaunchKernelGGL(grad_reduce_kernel, dim3((kParams + 255) / 256), dim3(256), 0, s, gpart, lpart, nblk, g, loss);
  HDG_HIP(hipGetLastError());
}

// Adam on the f32 master weights; writes the bf16-rounded working copy.
__global__ void adam_kernel(float* __restrict__ master, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ g, float* __restrict__ wq, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kParams) return;
  const float gi = g[i];
  m[i] = b1 * m[i] + (1.0f - b1) * gi;
  v[i] = b2 * v[i] + (1.0f - b2) * gi * gi;
  master[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
  wq[i] = bf16r(master[i]);
}

__global__ void predict_kernel(const float* __restrict__ X, uint64_t n, const float* __restrict__ wq,
                               float* __restrict__ out) {
  __shared__ float w[kParams];
  for (int i = threadIdx.x; i < kParams; i += blockDim.x) w[i] = wq[i];
  __syncthreads();
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[kF];
  for (int i = 0; i < kF; ++i) x[i] = bf16r(X[s * kF + i]);
  out[s] = mlp_forward(w, x, nullptr, nullptr);
}

}  // namespace

struct Index::BudgetModel {
  bool ready = false;
  std::vector<uint32_t> ladT, ladS;
  std::vector<float> master, mean, inv_std;
  float margin = 0.0f;
  DevBuf d_w, d_mean, d_inv, d_ladT, d_ladS;
  // feature capture during a fit
  float* sink = nullptr;
  uint64_t sink_rows = 0;
  BudgetReport report;
};

bool Index::budget_model_ready() const { return budget_ && budget_->ready; }

bool Index::budget_capturing() const { return budget_ && budget_->sink; }

bool Index::budget_probes_all(uint32_t C) const {
  if (!budget_) return true;
  for (uint32_t t : budget_->ladT)
    if (t < C) return false;
  return true;
}

void Index::run_budget_model(uint64_t qn, int l, const float* ranges, const uint32_t* box_lo, const uint32_t* box_w,
                             int n_valid, const float* feat_dist, uint32_t* n_chosen_s, uint32_t* budget_T,
                             uint32_t* budget_S, hipStream_t s, const uint32_t* abort_flag) {
  (void)box_lo;
  BudgetModel& B = *budget_;
  hipLaunchKernelGGL(infer_kernel, dim3((unsigned)std::min<uint64_t>((qn + kInferQ - 1) / kInferQ, kInferWgs)), dim3(256), 0, s, qn, l, ranges, box_w,
                     n_valid, feat_dist, B.d_w.as<float>(), B.d_mean.as<float>(), B.d_inv.as<float>(), B.margin,
                     (int)B.ladT.size(), B.d_ladT.as<uint32_t>(), B.d_ladS.as<uint32_t>(), n_chosen_s, budget_T,
                     budget_S, abort_flag);
  HDG_HIP(hipGetLastError());
}

void Index::capture_features(uint64_t q0, uint64_t qn, int l, const float* ranges, const uint32_t* box_w,
                             int n_valid, const float* feat_dist, hipStream_t s) {
  if (!budget_ || !budget_->sink) return;
  if (q0 + qn > budget_->sink_rows) fail(kBadState, "feature capture beyond the training set");
  hipLaunchKernelGGL(features_kernel, dim3((unsigned)((qn + 255) / 256)), dim3(256), 0, s, qn, l, ranges, box_w,
                     n_valid, feat_dist, budget_->sink + q0 * kF);
  HDG_HIP(hipGetLastError());
}

// recall@k of every query on the device (one wave per query): distinct valid ground-truth ids (< 2^63) found among
// the result ids, over the distinct valid ground-truth ids (1 when there are none).  Pairwise scans (k^2 per
// query), which is nothing next to the query calls it scores.
__global__ __launch_bounds__(256) void recall_kernel(const uint64_t* __restrict__ res, const uint64_t* __restrict__ gt,
                                                     uint64_t n, uint64_t k, float* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const uint64_t q = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (q >= n) return;
  const uint64_t* a = res + q * k;
  const uint64_t* b = gt + q * k;
  uint32_t hit = 0, nb = 0;
  for (uint64_t j = (uint64_t)lane; j < k; j += 64) {
    const uint64_t x = b[j];
    bool first = x < (1ull << 63);
    for (uint64_t i = 0; i < j && first; ++i) first = b[i] != x;
    if (!first) continue;
    ++nb;  // a distinct valid ground-truth id: found among the results?
    bool found = false;
    for (uint64_t i = 0; i < k && !found; ++i) found = a[i] == x;
    hit += found ? 1u : 0u;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    hit += (uint32_t)__shfl_xor((int)hit, o, 64);
    nb += (uint32_t)__shfl_xor((int)nb, o, 64);
  }
  if (lane == 0) out[q] = nb ? (float)hit / (float)nb : 1.0f;
}

// recall_kernel over n queries: results on the device, ground truth on the device.
static std::vector<float> recall_on_device(const uint64_t* d_res, const uint64_t* d_gt, uint64_t n, uint64_t k,
                                           DevBuf& d_out, hipStream_t s) {
  d_out.reserve(n * 4);
  hipLaunchKernelGGL(recall_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, d_res, d_gt, n, k,
                     d_out.as<float>());
  HDG_HIP(hipGetLastError());
  std::vector<float> r(n);
  HDG_HIP(hipMemcpyAsync(r.data(), d_out.data(), n * 4, hipMemcpyDeviceToHost, s));
  HDG_HIP(hipStreamSynchronize(s));
  return r;
}

void Index::budget_fit(const float* queries, uint64_t n, uint64_t d, const float* qlabels, uint64_t qcols,
                       const uint64_t* gt_host, uint64_t k, const BudgetFitParams& P, hipStream_t s) {
  HDG_HIP(hipSetDevice(device_));
  if (!st_.built) fail(kBadState, "budget fit before build");
  if (k != cfg_.topk) fail(kInvalidArgument, "ground truth width differs from TOPK");
  if (P.levels < 2 || P.levels > 32) fail(kInvalidArgument, "levels must be in [2, 32]");
  // the budget is applied after the sub-cluster scoring, whose scores are the features: every level probes
  // all non-zero clusters and only the sub-cluster budget grows
  if (P.t0 > 0) fail(kUnsupported, "cluster budgets (t0 > 0) are not supported; the model budgets sub-clusters");
  if (!budget_) budget_ = new BudgetModel();
  BudgetModel& B = *budget_;
  B.ready = false;
  // phase wall times at log level 1 (HDG_LOG=1)
  auto t_last = std::chrono::steady_clock::now();
  auto phase_done = [&](const char* what) {
    if (log_level() < 1) return;
    HDG_HIP(hipStreamSynchronize(s));
    const auto now = std::chrono::steady_clock::now();
    log_line(1, "budget fit %s: %.1f ms", what, std::chrono::duration<double, std::milli>(now - t_last).count());
    t_last = now;
  };
  // ladder
  const uint32_t n_list = (uint32_t)std::min<uint64_t>(cfg_.n_list, st_.C);
  const uint32_t sub_lists = (uint32_t)cfg_.sub_lists;
  B.ladT.resize(P.levels);
  B.ladS.resize(P.levels);
  for (int i = 0; i < P.levels; ++i) {
    const double f = std::pow(P.ratio, i);
    // t0 <= 0: every level probes all non-zero clusters (full counts) and only the sub-cluster budget grows
    B.ladT[i] = P.t0 > 0 ? (uint32_t)std::min<double>(n_list, std::max(1.0, std::round(P.t0 * f))) : n_list;
    B.ladS[i] = (uint32_t)std::min<double>(sub_lists, std::max(1.0, std::round(P.s0 * f)));
  }
  // traces: recall of every training query at every level of the labelling path (BUDGET_LABELS 0: the plain
  // reference-exact path, the minimal plain budget of each query; 1: the mixed path), then one mixed pass at the
  // top level that writes the feature rows
  const Config saved = cfg_;
  cfg_.precision = cfg_.budget_labels ? 1 : 0;
  cfg_.budget_model = 0;
  cfg_.graph = 0;
  DevBuf d_dist, d_ids, d_feat;
  d_dist.reserve(n * k * 4);
  d_ids.reserve(n * k * 8);
  d_feat.reserve(n * kF * 4);
  std::vector<std::vector<float>> rec(P.levels);
  DevBuf d_gt, d_rec;
  d_gt.reserve(n * k * 8);
  HDG_HIP(hipMemcpyAsync(d_gt.data(), gt_host, n * k * 8, hipMemcpyHostToDevice, s));
  HDG_HIP(hipStreamSynchronize(s));
  try {
    for (int i = 0; i < P.levels; ++i) {
      cfg_.probe_clusters = P.t0 > 0 ? B.ladT[i] : 0;
      cfg_.sub_lists = B.ladS[i];
      query(queries, n, d, qlabels, qcols, d_dist.as<float>(), d_ids.as<uint64_t>(), s);
      rec[i] = recall_on_device(d_ids.as<uint64_t>(), d_gt.as<uint64_t>(), n, k, d_rec, s);
    }
    cfg_.precision = 1;
    B.sink = d_feat.as<float>();
    B.sink_rows = n;
    query(queries, n, d, qlabels, qcols, d_dist.as<float>(), d_ids.as<uint64_t>(), s);
    B.sink = nullptr;
  } catch (...) {
    cfg_ = saved;
    B.sink = nullptr;
    throw;
  }
  cfg_ = saved;
  phase_done("labelling runs and feature capture");
  // labels: smallest level reaching the target recall (the top level when none does)
  std::vector<float> y(n);
  for (uint64_t q = 0; q < n; ++q) {
    int lev = P.levels - 1;
    for (int i = 0; i < P.levels; ++i)
      if (rec[i][q] >= P.target) {
        lev = i;
        break;
      }
    y[q] = (float)lev;
  }
  std::vector<float> X(n * kF);
  HDG_HIP(hipMemcpy(X.data(), d_feat.data(), n * kF * 4, hipMemcpyDeviceToHost));
  // split: the last holdout share of a seeded permutation calibrates the margin
  std::vector<uint32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0u);
  std::mt19937_64 rng(P.seed);
  std::shuffle(perm.begin(), perm.end(), rng);
  const uint64_t n_hold = std::max<uint64_t>(1, (uint64_t)(n * P.holdout));
  const uint64_t n_train = n - n_hold;
  B.mean.assign(kF, 0.0f);
  B.inv_std.assign(kF, 1.0f);
  for (int f = 0; f < kF; ++f) {
    double s1 = 0, s2 = 0;
    for (uint64_t t = 0; t < n_train; ++t) {
      double v = X[(uint64_t)perm[t] * kF + f];
      s1 += v;
      s2 += v * v;
    }
    double mu = s1 / n_train, var = s2 / n_train - mu * mu;
    B.mean[f] = (float)mu;
    B.inv_std[f] = var > 1e-12 ? (float)(1.0 / std::sqrt(var)) : 1.0f;
  }
  for (uint64_t q = 0; q < n; ++q)
    for (int f = 0; f < kF; ++f) X[q * kF + f] = (X[q * kF + f] - B.mean[f]) * B.inv_std[f];
  phase_done("labels and standardisation");
  // init (He) and train on the GPU
  B.master.assign(kParams, 0.0f);
  {
    std::normal_distribution<float> N01(0.0f, 1.0f);
    for (int i = 0; i < kP1; ++i) B.master[kOffW1 + i] = N01(rng) * std::sqrt(2.0f / kF);
    for (int i = 0; i < kP2; ++i) B.master[kOffW2 + i] = N01(rng) * std::sqrt(2.0f / kH);
    for (int i = 0; i < kH; ++i) B.master[kOffW3 + i] = N01(rng) * std::sqrt(1.0f / kH);
    float ymean = 0;
    for (uint64_t t = 0; t < n_train; ++t) ymean += y[perm[t]];
    B.master[kOffB3] = ymean / (float)n_train;
  }
  DevBuf dX, dY, dIdx, dMaster, dM, dV, dG, dWq, dLoss, dWork;
  upload(dX, X, s);
  upload(dY, y, s);
  upload(dMaster, B.master, s);
  std::vector<float> zeros(kParams, 0.0f), wq(kParams);
  upload(dM, zeros, s);
  upload(dV, zeros, s);
  dG.reserve(kParams * 4);
  for (int i = 0; i < kParams; ++i) {
    uint32_t u;
    std::memcpy(&u, &B.master[i], 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    std::memcpy(&wq[i], &u, 4);
  }
  upload(dWq, wq, s);
  dLoss.reserve(4);
  dIdx.reserve(n_train * 4);
  std::vector<uint32_t> order(perm.begin(), perm.begin() + n_train);
  const uint32_t bsz = (uint32_t)std::min<uint64_t>(P.batch, n_train);
  dWork.reserve(train_step_workspace_floats(bsz) * 4);
  int step = 0;
  float last_loss = 0.0f;
  for (int ep = 0; ep < P.epochs; ++ep) {
    std::shuffle(order.begin(), order.end(), rng);
    upload(dIdx, order, s);
    int nb = 0;
    HDG_HIP(hipMemsetAsync(dLoss.data(), 0, 4, s));  // the epoch's batch losses add up here
    for (uint64_t b0 = 0; b0 + bsz <= n_train; b0 += bsz) {
      launch_train_step(dX.as<float>(), dY.as<float>(), dIdx.as<uint32_t>() + b0, bsz, dWq.as<float>(), dG.as<float>(),
                        dLoss.as<float>(), dWork.as<float>(), s);
      ++step;
      const float c1 = 1.0f - std::pow(0.9f, (float)step), c2 = 1.0f - std::pow(0.999f, (float)step);
      hipLaunchKernelGGL(adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, s, dMaster.as<float>(),
                         dM.as<float>(), dV.as<float>(), dG.as<float>(), dWq.as<float>(), P.lr, 0.9f, 0.999f, 1e-8f,
                         c1, c2);
      HDG_HIP(hipGetLastError());
      ++nb;
    }
    float lv = 0.0f;
    HDG_HIP(hipMemcpyAsync(&lv, dLoss.data(), 4, hipMemcpyDeviceToHost, s));
    HDG_HIP(hipStreamSynchronize(s));  // also orders the next epoch's index upload after this epoch's steps
    last_loss = nb ? lv / nb : 0.0f;
  }
  HDG_HIP(hipMemcpy(B.master.data(), dMaster.data(), kParams * 4, hipMemcpyDeviceToHost));
  HDG_HIP(hipMemcpy(wq.data(), dWq.data(), kParams * 4, hipMemcpyDeviceToHost));
  phase_done("training");
  // predictions for all samples (mean level of the holdout at the chosen margin)
  DevBuf dPred;
  dPred.reserve(n * 4);
  hipLaunchKernelGGL(predict_kernel, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, dX.as<float>(), n,
                     dWq.as<float>(), dPred.as<float>());
  HDG_HIP(hipGetLastError());
  std::vector<float> pred(n);
  HDG_HIP(hipMemcpy(pred.data(), dPred.data(), n * 4, hipMemcpyDeviceToHost));
  auto level_of = [&](float yhat, float margin) {
    if (!std::isfinite(yhat)) return P.levels - 1;
    float t = std::ceil(yhat + margin);
    return t < 0 ? 0 : (t > P.levels - 1 ? P.levels - 1 : (int)t);
  };
  upload(B.d_w, wq, s);
  upload(B.d_mean, B.mean, s);
  upload(B.d_inv, B.inv_std, s);
  upload(B.d_ladT, B.ladT, s);
  upload(B.d_ladS, B.ladS, s);
  drop_graphs();  // captured graphs hold the previous model's margin and level count
  // margin: the smallest one (negative ones included, steps of 1/8 level) at which the deployed path -- mixed
  // precision with this model's budgets -- reaches the target mean recall on the holdout queries, run end to
  // end against the caller's ground truth
  const uint64_t n_hold_rows = n - n_train;
  std::vector<float> hq(n_hold_rows * d), hql(n_hold_rows * qcols), allq(n * d), allql(n * qcols);
  HDG_HIP(hipMemcpy(allq.data(), queries, n * d * 4, hipMemcpyDeviceToHost));
  HDG_HIP(hipMemcpy(allql.data(), qlabels, n * qcols * 4, hipMemcpyDeviceToHost));
  std::vector<uint64_t> hgt(n_hold_rows * k);
  for (uint64_t t = 0; t < n_hold_rows; ++t) {
    const uint64_t q = perm[n_train + t];
    std::memcpy(&hq[t * d], &allq[q * d], d * 4);
    std::memcpy(&hql[t * qcols], &allql[q * qcols], qcols * 4);
    std::memcpy(&hgt[t * k], &gt_host[q * k], k * 8);
  }
  DevBuf dhq, dhql, dhgt;
  upload(dhq, hq, s);
  upload(dhql, hql, s);
  upload(dhgt, hgt, s);
  float chosen = P.max_margin;
  double hold_rec = 0.0, hold_lev = 0.0;
  cfg_.precision = 1;
  cfg_.budget_model = 1;
  cfg_.graph = 0;
  cfg_.probe_clusters = 0;
  B.ready = true;
  try {
    for (float m = -P.max_margin; m <= P.max_margin + 1e-6f; m += 0.125f) {
      B.margin = m;
      query(dhq.as<float>(), n_hold_rows, d, dhql.as<float>(), qcols, d_dist.as<float>(), d_ids.as<uint64_t>(), s);
      const std::vector<float> r = recall_on_device(d_ids.as<uint64_t>(), dhgt.as<uint64_t>(), n_hold_rows, k, d_rec, s);
      double rsum = 0, rsq = 0;
      for (float v : r) {
        rsum += v;
        rsq += (double)v * v;
      }
      hold_rec = rsum / n_hold_rows;
      // the holdout mean must clear the target by one standard error, so that the recall on unseen queries of
      // the same distribution meets it more often than not despite the holdout's sampling noise
      const double var = std::max(0.0, rsq / n_hold_rows - hold_rec * hold_rec);
      const double se = std::sqrt(var / (double)n_hold_rows);
      chosen = m;
      if (hold_rec >= (double)P.target + se) break;
    }
  } catch (...) {
    cfg_ = saved;
    B.ready = false;
    throw;
  }
  cfg_ = saved;
  phase_done("margin calibration");
  B.margin = chosen;
  for (uint64_t t = n_train; t < n; ++t) hold_lev += level_of(pred[perm[t]], chosen);
  hold_lev /= (double)n_hold_rows;
  double mean_label = 0, fixed_rec = 0;
  for (uint64_t q = 0; q < n; ++q) {
    mean_label += y[q];
    fixed_rec += rec[P.levels - 1][q];
  }
  B.report = BudgetReport{last_loss, B.margin, (float)hold_rec, (float)hold_lev, (float)(mean_label / n),
                          (float)(fixed_rec / n), (uint64_t)n_train, (uint64_t)n_hold_rows};
  drop_graphs();
  B.ready = true;
}

void Index::budget_debug_features(const float* queries, uint64_t n, uint64_t d, const float* qlabels,
                                  uint64_t qcols, float* out_host, hipStream_t s) {
  HDG_HIP(hipSetDevice(device_));
  if (!st_.built) fail(kBadState, "feature capture before build");
  if (!budget_) budget_ = new BudgetModel();
  BudgetModel& B = *budget_;
  DevBuf d_feat, d_dist, d_ids;
  d_feat.reserve(n * kF * 4);
  d_dist.reserve(n * cfg_.topk * 4);
  d_ids.reserve(n * cfg_.topk * 8);
  const Config saved = cfg_;
  cfg_.precision = 1;
  cfg_.budget_model = 0;
  cfg_.graph = 0;
  B.sink = d_feat.as<float>();
  B.sink_rows = n;
  try {
    query(queries, n, d, qlabels, qcols, d_dist.as<float>(), d_ids.as<uint64_t>(), s);
  } catch (...) {
    B.sink = nullptr;
    cfg_ = saved;
    throw;
  }
  B.sink = nullptr;
  cfg_ = saved;
  HDG_HIP(hipMemcpy(out_host, d_feat.data(), n * kF * 4, hipMemcpyDeviceToHost));
}

const BudgetReport& Index::budget_report() const {
  if (!budget_) fail(kBadState, "no budget model");
  return budget_->report;
}

void Index::budget_save(const std::string& path) const {
  if (!budget_model_ready()) fail(kBadState, "no budget model to save");
  const BudgetModel& B = *budget_;
  std::ofstream f(path, std::ios::binary);
  if (!f) fail(kIoError, "cannot write " + path);
  const uint32_t magic = 0x48444742u, levels = (uint32_t)B.ladT.size(), params = kParams;
  f.write((const char*)&magic, 4);
  f.write((const char*)&levels, 4);
  f.write((const char*)&params, 4);
  f.write((const char*)&B.margin, 4);
  f.write((const char*)B.ladT.data(), levels * 4);
  f.write((const char*)B.ladS.data(), levels * 4);
  f.write((const char*)B.mean.data(), kF * 4);
  f.write((const char*)B.inv_std.data(), kF * 4);
  f.write((const char*)B.master.data(), kParams * 4);
}

void Index::budget_load(const std::string& path, hipStream_t s) {
  HDG_HIP(hipSetDevice(device_));
  std::ifstream f(path, std::ios::binary);
  if (!f) fail(kIoError, "cannot read " + path);
  uint32_t magic = 0, levels = 0, params = 0;
  float margin = 0;
  f.read((char*)&magic, 4);
  f.read((char*)&levels, 4);
  f.read((char*)&params, 4);
  f.read((char*)&margin, 4);
  if (!f || magic != 0x48444742u || params != (uint32_t)kParams || levels < 2 || levels > 32)
    fail(kIoError, "not a budget model file: " + path);
  // the whole file is read and checked before anything of the current model changes
  std::vector<uint32_t> ladT(levels), ladS(levels);
  std::vector<float> mean(kF), inv_std(kF), master(kParams);
  f.read((char*)ladT.data(), levels * 4);
  f.read((char*)ladS.data(), levels * 4);
  f.read((char*)mean.data(), kF * 4);
  f.read((char*)inv_std.data(), kF * 4);
  f.read((char*)master.data(), kParams * 4);
  if (!f) fail(kIoError, "short budget model file: " + path);
  if (!std::isfinite(margin)) fail(kIoError, "budget model margin is not finite: " + path);
  for (uint32_t i = 0; i < levels; ++i)
    if (ladS[i] == 0 || ladT[i] == 0) fail(kIoError, "budget model ladder holds a zero budget: " + path);
  std::vector<float> wq(kParams);
  for (int i = 0; i < kParams; ++i) {
    uint32_t u;
    std::memcpy(&u, &master[i], 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    std::memcpy(&wq[i], &u, 4);
  }
  // captured graphs hold the old model's margin and level count as kernel arguments
  drop_graphs();
  if (!budget_) budget_ = new BudgetModel();
  BudgetModel& B = *budget_;
  B.ready = false;
  upload(B.d_w, wq, s);
  upload(B.d_mean, mean, s);
  upload(B.d_inv, inv_std, s);
  upload(B.d_ladT, ladT, s);
  upload(B.d_ladS, ladS, s);
  B.ladT = std::move(ladT);
  B.ladS = std::move(ladS);
  B.mean = std::move(mean);
  B.inv_std = std::move(inv_std);
  B.master = std::move(master);
  B.margin = margin;
  B.ready = true;
}

void Index::budget_free() {
  delete budget_;
  budget_ = nullptr;
}

// One forward/backward of the training kernel on a given batch (gradient check against a CPU oracle).
void budget_debug_grads(const float* params, const float* X, const float* Y, uint32_t n, float* grads, float* loss) {
  DevBuf dX, dY, dIdx, dW, dG, dL, dWork;
  std::vector<float> x(X, X + (size_t)n * kF), y(Y, Y + n), w(params, params + kParams);
  std::vector<uint32_t> idx(n);
  std::iota(idx.begin(), idx.end(), 0u);
  upload(dX, x, nullptr);
  upload(dY, y, nullptr);
  upload(dIdx, idx, nullptr);
  upload(dW, w, nullptr);
  dG.reserve(kParams * 4);
  dL.reserve(4);
  dWork.reserve(train_step_workspace_floats(n) * 4);
  HDG_HIP(hipMemset(dL.data(), 0, 4));
  launch_train_step(dX.as<float>(), dY.as<float>(), dIdx.as<uint32_t>(), n, dW.as<float>(), dG.as<float>(),
                    dL.as<float>(), dWork.as<float>(), nullptr);
  HDG_HIP(hipDeviceSynchronize());
  HDG_HIP(hipMemcpy(grads, dG.data(), kParams * 4, hipMemcpyDeviceToHost));
  HDG_HIP(hipMemcpy(loss, dL.data(), 4, hipMemcpyDeviceToHost));
}

int budget_param_count() { return kParams; }

}  // namespace hdg
