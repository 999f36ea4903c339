// Launchers of the mixed-precision, budgeted front half of the query path (hdg_fast.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hdg {

constexpr int kFeatDists = 32;  // coarse distances kept per query as budget-model features

// queries[q0 .. q0+qn) [d] f32 -> out [qn][Dp] bf16 (round to nearest even, zero padded).
void launch_to_bf16(const float* queries, uint64_t q0, uint64_t qn, uint64_t d, uint64_t Dp, uint16_t* out,
                    hipStream_t s);
// S[q][c] = cnorm[c] - 2 <q, c> (BF16 MFMA, FP32 accumulate); qb16 = queries as bf16 [qn][Dp], cb16 =
// centers as bf16 [C][Dp], Dp % 32 == 0.  With nzb/nzp (per query nw bit words of its non-zero items and
// the rank of each word's first item) row q keeps only its non-zero items, item c at rank
// nzp[q][c/32] + popcount(nzb[q][c/32] below bit c%32); rows have stride C either way.  abort_flag: the
// kernel returns at once when it is set.
void launch_coarse_gemm(const uint16_t* qb16, uint64_t qn, uint64_t Dp, const uint16_t* cb16, const float* cnorm,
                        uint32_t C, float* S, hipStream_t s, const uint32_t* nzb = nullptr,
                        const uint32_t* nzp = nullptr, uint32_t nw = 0, const uint32_t* abort_flag = nullptr);
// Budget-model features of queries whose every non-zero cluster is probed: feat [qn][kFeatDists] = the 32
// smallest of S[q][list[q][0 .. n_list_q[q])] ascending (the last repeated when fewer), each + |q|^2; the
// values launch_scored_select writes as features.  cap (the list row stride) <= 512.
void launch_coarse_features(const float* queries, uint64_t q0, uint64_t qn, uint64_t d, const uint32_t* list,
                            uint32_t cap, const uint32_t* n_list_q, const float* S, uint64_t ld, float* feat,
                            hipStream_t s, const uint32_t* abort_flag);
// Per query: top-k (k = budget[q] or k_default) of the listed items list[q][0..n_list_q[q]) by S[q][item],
// ascending; writes the chosen list positions chosen[q][0..m) (row stride out_ld) and n_chosen[q];
// optional features.
// scratch: [qn][cap] u64, used when a list exceeds the LDS sort capacity.  wave: lists of at most 512 items
// are sorted by one wave per query in registers (same result as the workgroup sort).  nzb/nzp: S holds
// compact rows (launch_coarse_gemm) and every listed item is one of the query's non-zero items.
// abort_flag: every kernel returns at once when it is set (graph replay); `left` counts the queries the wave
// kernel leaves to the workgroup kernel; skip_fallback (with abort_flag) launches no workgroup kernel and
// sets the flag instead when a query needs it.  by_pos: the list is exactly the query's non-zero items in
// ascending order, so item e's compact score sits at row position e (no rank lookup).
void launch_scored_select(const float* queries, uint64_t q0, uint64_t qn, uint64_t d, const uint32_t* list,
                          uint32_t cap, const uint32_t* n_list_q, const float* S, uint64_t ld, uint32_t k_default,
                          const uint32_t* budget, uint32_t out_ld, uint32_t* chosen, uint32_t* n_chosen,
                          float* feat_dist, uint64_t* scratch, bool wave, hipStream_t s,
                          const uint32_t* nzb = nullptr, const uint32_t* nzp = nullptr, uint32_t nw = 0,
                          uint32_t* abort_flag = nullptr, uint32_t* left = nullptr, bool skip_fallback = false,
                          bool by_pos = false);
}  // namespace hdg
