// Mixed-precision, budgeted front half of the query path (PRECISION = 1).
//
// The plain path (hdg_query.hip) reproduces the reference exactly: it counts every sub-cluster for every
// query and scores the non-zero clusters with exact FP32 L2.  This file replaces that front half with
//   1. a BF16 MFMA contraction queries x cluster centers (FP32 accumulate): score = |c|^2 - 2 q.c;
//   2. a per-query top-T of those scores (T = the query's probe budget, fixed or predicted);
//   3. exact in-box counts for the sub-clusters of those T clusters only.
// Everything after (non-zero sub-clusters, exact FP32 sub-center scoring, Morton ranges, PQ, label filter,
// exact FP32 re-rank) runs on the same kernels as the plain path.  Results are approximate in the coarse
// choice only; recall is measured against the FP64 brute force.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "hdg_device.hpp"
#include "hdg_fast.hpp"

namespace hdg {

namespace {

using dev::kWG;
typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ short f32_to_bf16(float v) {
  uint32_t u = __float_as_uint(v);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (inputs are finite)
  return (short)(u >> 16);
}

// S[q][c] = cnorm[c] - 2 * <q, c> with bf16 operands and f32 accumulation.  Block = 4 waves = 64 queries
// x 64 centers; each wave owns 16 queries x 64 centers as four 16x16 MFMA accumulators and walks K in
// steps of 32 (v_mfma_f32_16x16x32_bf16: lane l supplies A[l%16][8*(l/16)..+7] and B[8*(l/16)..+7][l%16],
// and receives D[4*(l/16)..+3][l%16]).
__global__ __launch_bounds__(256) void coarse_gemm_kernel(const float* __restrict__ queries, uint64_t q0,
                                                          uint64_t qn, uint64_t d, uint64_t Dp,
                                                          const uint16_t* __restrict__ cb16,
                                                          const float* __restrict__ cnorm, uint32_t C,
                                                          float* __restrict__ S) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint64_t qb = (uint64_t)blockIdx.x * 64 + wave * 16;
  const uint32_t cbase = blockIdx.y * 64;
  const int r = lane & 15, kg = lane >> 4;
  const uint64_t qrow = qb + r;
  const bool qok = qrow < qn;
  const float* qp = queries + (q0 + (qok ? qrow : 0)) * d;
  f32x4 acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
  const bf16x8 zero = {0, 0, 0, 0, 0, 0, 0, 0};
  for (uint64_t k0 = 0; k0 < Dp; k0 += 32) {
    bf16x8 a;
    const uint64_t kk = k0 + kg * 8;
#pragma unroll
    for (int t = 0; t < 8; ++t) a[t] = f32_to_bf16((qok && kk + t < d) ? qp[kk + t] : 0.0f);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t c = cbase + j * 16 + r;
      const bf16x8 b = c < C ? *reinterpret_cast<const bf16x8*>(cb16 + (uint64_t)c * Dp + kk) : zero;
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t c = cbase + j * 16 + (lane & 15);
    if (c >= C) continue;
    const float cn = cnorm[c];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint64_t q = qb + (lane >> 4) * 4 + i;
      if (q < qn) S[q * C + c] = cn - 2.0f * acc[j][i];
    }
  }
}

// order-preserving u32 key of a signed f32
__device__ __forceinline__ uint32_t ord(float v) {
  uint32_t u = __float_as_uint(v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Per query: the T = min(budget, n_list) clusters with the smallest coarse score, ascending, as the chosen
// clusters; also the query's squared norm and its 32 smallest coarse distances (budget-model features).
__global__ __launch_bounds__(256) void coarse_topk_kernel(const float* __restrict__ queries, uint64_t q0,
                                                          uint64_t d, const float* __restrict__ S, uint32_t C,
                                                          uint32_t n_list, uint32_t probe_default,
                                                          const uint32_t* __restrict__ budget_T,
                                                          uint32_t* __restrict__ chosen_c,
                                                          uint32_t* __restrict__ n_chosen_c,
                                                          float* __restrict__ feat_dist) {
  __shared__ uint64_t keys[dev::kSortCap];
  __shared__ float red[kWG];
  const uint64_t qi = blockIdx.x;
  const float* row = S + qi * C;
  for (uint32_t c = threadIdx.x; c < C; c += kWG) keys[c] = ((uint64_t)ord(row[c]) << 32) | c;
  // |q|^2 in f32 (feature offset)
  const float* q = queries + (q0 + qi) * d;
  float part = 0.0f;
  for (uint64_t t = threadIdx.x; t < d; t += kWG) part += q[t] * q[t];
  red[threadIdx.x] = part;
  __syncthreads();
  dev::wg_sort_lds(keys, C);
  if (threadIdx.x == 0) {
    float qq = 0.0f;
    for (int i = 0; i < kWG; ++i) qq += red[i];
    red[0] = qq;
  }
  __syncthreads();
  uint32_t T = budget_T ? budget_T[qi] : probe_default;
  if (T == 0 || T > n_list) T = n_list;
  if (T > C) T = C;
  for (uint32_t j = threadIdx.x; j < T; j += kWG) chosen_c[qi * n_list + j] = (uint32_t)keys[j];
  if (feat_dist) {
    for (uint32_t j = threadIdx.x; j < kFeatDists; j += kWG) {
      float v = 0.0f;
      if (j < C) {
        uint32_t o = (uint32_t)(keys[j] >> 32);
        uint32_t u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
        v = __uint_as_float(u) + red[0];
      }
      feat_dist[qi * kFeatDists + j] = v;
    }
  }
  if (threadIdx.x == 0) n_chosen_c[qi] = T;
}

// Exact in-box counts for the sub-clusters of each query's chosen clusters (same integers as the plain
// count); one workgroup per query, one thread per sub-cluster, Morton range by binary search in HBM.
template <int NV>
__global__ __launch_bounds__(256) void budget_count_kernel(const uint32_t* __restrict__ cells, uint64_t N,
                                                           const uint64_t* __restrict__ zcode,
                                                           const uint32_t* __restrict__ seg_off,
                                                           const uint32_t* __restrict__ sub_off, uint64_t CS,
                                                           const uint32_t* __restrict__ box_lo,
                                                           const uint32_t* __restrict__ box_w,
                                                           const uint64_t* __restrict__ zlo_cnt,
                                                           const uint64_t* __restrict__ zhi, uint64_t qn,
                                                           uint32_t n_list, const uint32_t* __restrict__ chosen_c,
                                                           const uint32_t* __restrict__ n_chosen_c,
                                                           uint32_t* __restrict__ counts) {
  __shared__ uint32_t pre[1025];
  const uint64_t qi = blockIdx.x;
  const uint32_t nc = n_chosen_c[qi];
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (uint32_t j = 0; j < nc; ++j) {
      pre[j] = acc;
      const uint32_t c = chosen_c[qi * n_list + j];
      acc += sub_off[c + 1] - sub_off[c];
    }
    pre[nc] = acc;
  }
  __syncthreads();
  uint32_t lo[NV], w[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    lo[j] = box_lo[(uint64_t)j * qn + qi];
    w[j] = box_w[(uint64_t)j * qn + qi];
  }
  const uint64_t za = zlo_cnt[qi], zb = zhi[qi];
  const uint32_t total = pre[nc];
  for (uint32_t t = threadIdx.x; t < total; t += kWG) {
    uint32_t a = 0, b = nc;
    while (b - a > 1) {
      const uint32_t mid = (a + b) >> 1;
      if (pre[mid] <= t) a = mid; else b = mid;
    }
    const uint32_t sid = sub_off[chosen_c[qi * n_list + a]] + (t - pre[a]);
    uint32_t p0 = seg_off[sid], p1 = seg_off[sid + 1];
    uint32_t x = p0, y = p1;
    while (x < y) {
      const uint32_t mid = (x + y) >> 1;
      if (zcode[mid] < za) x = mid + 1; else y = mid;
    }
    uint32_t u = x, v = p1;
    while (u < v) {
      const uint32_t mid = (u + v) >> 1;
      if (zcode[mid] <= zb) u = mid + 1; else v = mid;
    }
    uint32_t cnt = 0;
    for (uint32_t p = x; p < u; ++p) {
      bool in = true;
#pragma unroll
      for (int j = 0; j < NV; ++j) in = in && (cells[(uint64_t)j * N + p] - lo[j] <= w[j]);
      cnt += in ? 1u : 0u;
    }
    counts[qi * CS + sid] = cnt;
  }
}

// Top-k of a query's listed items by their row of scores (ascending, ties to the lower list position):
// the mixed-precision replacement of the exact-FP32 refine + select_k used for cluster (A5) and
// sub-cluster (A6) selection.  k = budget[q] when given (0 = k_default), clamped to k_default.
// Optionally writes the 32 smallest listed scores (+ |q|^2) as budget-model features.
__global__ __launch_bounds__(256) void scored_select_kernel(const float* __restrict__ queries, uint64_t q0,
                                                            uint64_t d, const uint32_t* __restrict__ list,
                                                            uint32_t cap, const uint32_t* __restrict__ n_list_q,
                                                            const float* __restrict__ S, uint64_t ld,
                                                            uint32_t k_default, const uint32_t* __restrict__ budget,
                                                            uint32_t out_ld, uint32_t* __restrict__ chosen,
                                                            uint32_t* __restrict__ n_chosen,
                                                            float* __restrict__ feat_dist,
                                                            uint64_t* __restrict__ scratch) {
  __shared__ uint64_t keys[dev::kSortCap];
  __shared__ uint32_t hist[2048];
  __shared__ uint64_t sh[8];
  __shared__ float red[kWG];
  const uint64_t qi = blockIdx.x;
  const uint32_t n = n_list_q[qi];
  const uint32_t* lst = list + qi * cap;
  const float* row = S + qi * ld;
  uint32_t k = k_default;
  if (budget) {
    const uint32_t b = budget[qi];
    if (b != 0 && b < k) k = b;
  }
  uint32_t m = n < k ? n : k;
  if (n <= (uint32_t)dev::kSortCap) {
    for (uint32_t j = threadIdx.x; j < n; j += kWG) keys[j] = ((uint64_t)ord(row[lst[j]]) << 32) | j;
    __syncthreads();
    dev::wg_sort_lds(keys, n);
  } else {
    uint64_t* g = scratch + qi * cap;
    for (uint32_t j = threadIdx.x; j < n; j += kWG) g[j] = ((uint64_t)ord(row[lst[j]]) << 32) | j;
    __syncthreads();
    if (m > (uint32_t)dev::kSortCap) m = dev::kSortCap;
    dev::wg_radix_select(g, n, m, keys, hist, sh);
    dev::wg_sort_lds(keys, m);
  }
  for (uint32_t j = threadIdx.x; j < m; j += kWG) chosen[qi * out_ld + j] = lst[(uint32_t)keys[j]];
  if (threadIdx.x == 0) n_chosen[qi] = m;
  if (feat_dist) {
    const float* q = queries + (q0 + qi) * d;
    float part = 0.0f;
    for (uint64_t t = threadIdx.x; t < d; t += kWG) part += q[t] * q[t];
    red[threadIdx.x] = part;
    __syncthreads();
    if (threadIdx.x == 0) {
      float qq = 0.0f;
      for (int i = 0; i < kWG; ++i) qq += red[i];
      red[0] = qq;
    }
    __syncthreads();
    const uint32_t avail = n < 32 ? n : 32;
    for (uint32_t j = threadIdx.x; j < (uint32_t)kFeatDists; j += kWG) {
      float v = 0.0f;
      if (avail > 0) {
        const uint32_t o = (uint32_t)(keys[j < avail ? j : avail - 1] >> 32);
        const uint32_t u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
        v = __uint_as_float(u) + red[0];
      }
      feat_dist[qi * kFeatDists + j] = v;
    }
  }
}

}  // namespace

void launch_scored_select(const float* queries, uint64_t q0, uint64_t qn, uint64_t d, const uint32_t* list,
                          uint32_t cap, const uint32_t* n_list_q, const float* S, uint64_t ld, uint32_t k_default,
                          const uint32_t* budget, uint32_t out_ld, uint32_t* chosen, uint32_t* n_chosen,
                          float* feat_dist, uint64_t* scratch, hipStream_t s) {
  hipLaunchKernelGGL(scored_select_kernel, dim3((unsigned)qn), dim3(256), 0, s, queries, q0, d, list, cap, n_list_q, S,
                     ld, k_default, budget, out_ld, chosen, n_chosen, feat_dist, scratch);
}

void launch_coarse_gemm(const float* queries, uint64_t q0, uint64_t qn, uint64_t d, uint64_t Dp,
                        const uint16_t* cb16, const float* cnorm, uint32_t C, float* S, hipStream_t s) {
  dim3 grid((unsigned)((qn + 63) / 64), (unsigned)((C + 63) / 64));
  hipLaunchKernelGGL(coarse_gemm_kernel, grid, dim3(256), 0, s, queries, q0, qn, d, Dp, cb16, cnorm, C, S);
}

void launch_coarse_topk(const float* queries, uint64_t q0, uint64_t qn, uint64_t d, const float* S, uint32_t C,
                        uint32_t n_list, uint32_t probe_default, const uint32_t* budget_T, uint32_t* chosen_c,
                        uint32_t* n_chosen_c, float* feat_dist, hipStream_t s) {
  hipLaunchKernelGGL(coarse_topk_kernel, dim3((unsigned)qn), dim3(256), 0, s, queries, q0, d, S, C, n_list,
                     probe_default, budget_T, chosen_c, n_chosen_c, feat_dist);
}

void launch_budget_count(int n_valid, const uint32_t* cells, uint64_t N, const uint64_t* zcode,
                         const uint32_t* seg_off, const uint32_t* sub_off, uint64_t CS, const uint32_t* box_lo,
                         const uint32_t* box_w, const uint64_t* zlo_cnt, const uint64_t* zhi, uint64_t qn,
                         uint32_t n_list, const uint32_t* chosen_c, const uint32_t* n_chosen_c, uint32_t* counts,
                         hipStream_t s) {
  switch (n_valid) {
#define HDG_BC(NV)                                                                                          \
  case NV:                                                                                                  \
    hipLaunchKernelGGL(budget_count_kernel<NV>, dim3((unsigned)qn), dim3(256), 0, s, cells, N, zcode, seg_off, \
                       sub_off, CS, box_lo, box_w, zlo_cnt, zhi, qn, n_list, chosen_c, n_chosen_c, counts);   \
    break;
    HDG_BC(1)
    HDG_BC(2)
    HDG_BC(3)
    HDG_BC(4)
    HDG_BC(5)
    HDG_BC(6)
    HDG_BC(7)
    HDG_BC(8)
#undef HDG_BC
    default:
      break;
  }
}

}  // namespace hdg
