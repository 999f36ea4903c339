// Launchers of the mixed-precision, budgeted front half of the query path (hdg_fast.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hdg {

constexpr int kFeatDists = 32;  // coarse distances kept per query as budget-model features

// S[q][c] = cnorm[c] - 2 <q, c> (BF16 MFMA, FP32 accumulate); cb16 = centers as bf16 [C][Dp], Dp % 32 == 0.
void launch_coarse_gemm(const float* queries, uint64_t q0, uint64_t qn, uint64_t d, uint64_t Dp,
                        const uint16_t* cb16, const float* cnorm, uint32_t C, float* S, hipStream_t s);
// Per-query top-T clusters by S (T from budget_T[q] when given, else probe_default; clamped to n_list).
void launch_coarse_topk(const float* queries, uint64_t q0, uint64_t qn, uint64_t d, const float* S, uint32_t C,
                        uint32_t n_list, uint32_t probe_default, const uint32_t* budget_T, uint32_t* chosen_c,
                        uint32_t* n_chosen_c, float* feat_dist, hipStream_t s);
// Per query: top-k (k = budget[q] or k_default) of the listed items list[q][0..n_list_q[q]) by S[q][item],
// ascending; writes chosen[q][0..m) (row stride out_ld) and n_chosen[q]; optional features.
// scratch: [qn][cap] u64, used when a list exceeds the LDS sort capacity.
void launch_scored_select(const float* queries, uint64_t q0, uint64_t qn, uint64_t d, const uint32_t* list,
                          uint32_t cap, const uint32_t* n_list_q, const float* S, uint64_t ld, uint32_t k_default,
                          const uint32_t* budget, uint32_t out_ld, uint32_t* chosen, uint32_t* n_chosen,
                          float* feat_dist, uint64_t* scratch, hipStream_t s);
// Exact in-box counts of the chosen clusters' sub-clusters, written to counts[q][sub].
void launch_budget_count(int n_valid, const uint32_t* cells, uint64_t N, const uint64_t* zcode,
                         const uint32_t* seg_off, const uint32_t* sub_off, uint64_t CS, const uint32_t* box_lo,
                         const uint32_t* box_w, const uint64_t* zlo_cnt, const uint64_t* zhi, uint64_t qn,
                         uint32_t n_list, const uint32_t* chosen_c, const uint32_t* n_chosen_c, uint32_t* counts,
                         hipStream_t s);

}  // namespace hdg
