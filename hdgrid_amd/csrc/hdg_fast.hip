// Mixed-precision cluster and sub-cluster choice (PRECISION = 1).
//
// The plain path (hdg_query.hip) scores the non-zero clusters and sub-clusters of every query with exact
// FP32 L2 in the reference's summation order.  This file replaces those two scorings with
//   1. a BF16 MFMA contraction queries x (sub-)centers with FP32 accumulation: score = |c|^2 - 2 q.c;
//   2. a per-query top-k of the listed items by that score (k = n_list / sub_lists, or the query's own
//      budget from the learned model).
// The in-box counts, Morton ranges, PQ, label filter and the exact FP32 re-rank run on the same kernels
// as the plain path, so results differ from it only through which clusters and sub-clusters are kept;
// recall is measured against the FP64 brute force.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "hdg_common.hpp"
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

// queries [qn][d] f32 -> bf16 [qn][Dp] (round to nearest even, zero padded).
__global__ void to_bf16_kernel(const float* __restrict__ queries, uint64_t q0, uint64_t qn, uint64_t d,
                               uint64_t Dp, uint16_t* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= qn * Dp) return;
  const uint64_t q = i / Dp, t = i % Dp;
  out[i] = (uint16_t)f32_to_bf16(t < d ? queries[(q0 + q) * d + t] : 0.0f);
}

// S[q][c] = cnorm[c] - 2 * <q, c> with bf16 operands and f32 accumulation.  Block = 8 waves = 128 queries
// x 128 centers; wave (wm, wn) (4 x 2) owns 32 x 64 as 2 x 4 accumulators of v_mfma_f32_16x16x32_bf16 (lane l
// supplies A[l%16][8*(l/16)..+7] and B[8*(l/16)..+7][l%16], receives D[4*(l/16)..+3][l%16]).  K advances in
// steps of 32.  The operand tiles of a step go from global memory straight into LDS (global_load_lds, 16 bytes
// per lane) through a ring of three step buffers, so the loads of the next two steps are in flight while a
// step's MFMAs run, without holding them in registers.  A step buffer is 16-byte pieces: row r's piece p
// (bf16 k0 + 8p .. +7) at slot 4r + (p ^ swz(r)) -- the rows have no padding and the XOR swizzle keeps the
// fragment reads (16 rows x one piece per 16 lanes) on distinct banks.  The center norms are fetched before
// the main loop, so the epilogue does not wait on them.  (Measured on 10k x 5,120 x 384 with 9% non-zero
// compact rows: 256 x 128 tiles of 64 x 64 waves took 80 us, these 68 us; the main loop alone is ~50 us at
// any tile shape tried, the rest is epilogue latency.)
//
// Tile order: workgroup w runs on XCD w % 8 as that XCD's (w / 8)-th workgroup.  w is mapped to a position
// p = (w % 8) * per + w / 8 of a grouped order that takes kGemmGroup query tiles for each center tile, so the
// workgroups resident on one XCD at a time cover a few query tiles x a dozen center tiles and find their
// operand tiles in that XCD's L2.
constexpr int kGemmTile = 128;   // centers per tile
constexpr int kGemmTileQ = 128;  // queries per tile: 8 waves as 4 x 2, 32 x 64 each
constexpr int kGemmThreads = 512;
constexpr int kGemmStages = 3;   // step buffers in the LDS ring
constexpr int kGemmStepPieces = (kGemmTileQ + kGemmTile) * 4;  // 16-byte pieces of one K step
constexpr uint32_t kGemmGroup = 4;
constexpr uint32_t kXcds = 8;

__device__ __forceinline__ int gemm_swz(int r) { return ((r >> 1) & 3) ^ ((r >> 3) & 1); }

typedef __attribute__((address_space(3))) void lds_void;
typedef const __attribute__((address_space(1))) void global_cvoid;


__global__ __launch_bounds__(512) void coarse_gemm_kernel(const uint16_t* __restrict__ qb16, uint64_t qn, uint64_t Dp,
                                                          const uint16_t* __restrict__ cb16,
                                                          const float* __restrict__ cnorm, uint32_t C,
                                                          float* __restrict__ S, uint32_t QT, uint32_t CT,
                                                          uint32_t per, const uint32_t* __restrict__ nzb,
                                                          const uint32_t* __restrict__ nzp, uint32_t nw,
                                                          uint64_t ld, const uint32_t* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  const uint32_t p = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
  if (p >= QT * CT) return;
  const uint32_t span = kGemmGroup * CT, first = (p / span) * kGemmGroup;
  const uint32_t gsize = QT - first < kGemmGroup ? QT - first : kGemmGroup;
  const uint32_t inner = p % span;
  // the step ring, reused by the epilogue as 8 waves x 16 rows x 65 floats
  static_assert(kGemmStages * kGemmStepPieces * 16 >= 8 * 16 * 65 * 4, "epilogue does not fit the ring");
  static_assert(kGemmStepPieces == 2 * kGemmThreads && kGemmTileQ * 4 == kGemmThreads, "one A and one B piece per lane");
  __shared__ __align__(16) uint4 ring[kGemmStages * kGemmStepPieces];
  // compact rows: the tile's 128 rows x 4 bit words of non-zero items and their ranks
  __shared__ uint32_t nz_bits[kGemmTileQ * 4], nz_rank[kGemmTileQ * 4];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave >> 1, wn = wave & 1;
  const uint64_t qbase = (uint64_t)(first + inner % gsize) * kGemmTileQ;
  const uint64_t cbase = (uint64_t)(inner / gsize) * kGemmTile;
  // this lane's global sources: one A piece (slot wave*64 + lane) and one B piece (slot (wave+8)*64 + lane,
  // i.e. B slot wave*64 + lane); rows past the operands read the last row (their results are never stored)
  const uint16_t* srcA;
  const uint16_t* srcB;
  {
    const int sl = wave * 64 + lane, r = sl >> 2, pc = (sl & 3) ^ gemm_swz(r);
    const uint64_t q = qbase + r < qn ? qbase + r : qn - 1;
    srcA = qb16 + q * Dp + pc * 8;
    const uint64_t c = cbase + r < C ? cbase + r : C - 1;
    srcB = cb16 + c * Dp + pc * 8;
  }
  auto issue = [&](uint32_t ks) {
    uint4* buf = ring + (ks % kGemmStages) * kGemmStepPieces;
    const uint64_t k0 = (uint64_t)ks * 32;
    __builtin_amdgcn_global_load_lds((global_cvoid*)(srcA + k0), (lds_void*)(buf + wave * 64), 16, 0, 0);
    __builtin_amdgcn_global_load_lds((global_cvoid*)(srcB + k0), (lds_void*)(buf + kGemmTileQ * 4 + wave * 64), 16,
                                     0, 0);
  };
  f32x4 acc[2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
  // the compact-row words and the center norms are fetched now; the words are parked in LDS after the main loop
  const int r16 = lane & 15, kg = lane >> 4;
  uint32_t nzv[2] = {0u, 0u};
  if (nzb) {
    const int r = tid >> 2, wd = tid & 3;
    const uint64_t q = qbase + r;
    const uint64_t wi = (cbase >> 5) + (uint64_t)wd;
    if (q < qn && wi < nw) {
      nzv[0] = nzb[q * nw + wi];
      nzv[1] = nzp[q * nw + wi];
    }
  }
  float cn[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint64_t c = cbase + wn * 64 + j * 16 + r16;
    cn[j] = c < C ? cnorm[c] : 0.0f;
  }
  const uint32_t KS = (uint32_t)(Dp / 32);
  issue(0);
  if (KS > 1) issue(1);
  const int frag = r16 * 4 + (kg ^ gemm_swz(r16));  // slot of this lane's piece in a 16-row group
  const uint32_t ring_lds = (uint32_t)(uintptr_t)(lds_void*)ring;
  const uint32_t offA = (uint32_t)((wm * 32 * 4 + frag) * 16), offB = (uint32_t)((kGemmTileQ * 4 + wn * 64 * 4 + frag) * 16);
  for (uint32_t ks = 0; ks < KS; ++ks) {
    // step ks has landed once at most the next step's two loads of this lane are outstanding; the barrier
    // makes every lane's pieces visible and ends every wave's reads of step ks - 1, whose buffer step ks + 2
    // now refills.  Wait, barrier and fragment reads are written out (the compiler's fence and its alias
    // check against the LDS loads in flight would wait for every outstanding load instead).  The read block
    // waits for its first four fragments only; the MFMAs that need just those run while the last two B
    // fragments land, and the second wait takes those two as operands, so no use of them precedes it.
    if (ks + 1 < KS) asm volatile("s_waitcnt vmcnt(2)\n\ts_barrier" ::: "memory");
    else asm volatile("s_waitcnt vmcnt(0)\n\ts_barrier" ::: "memory");
    if (ks + 2 < KS) issue(ks + 2);
    const uint32_t bufb = ring_lds + (ks % kGemmStages) * (kGemmStepPieces * 16);
    bf16x8 a[2], b[4];
    asm volatile(
        "ds_read_b128 %0, %6\n\t"
        "ds_read_b128 %1, %6 offset:1024\n\t"
        "ds_read_b128 %2, %7\n\t"
        "ds_read_b128 %3, %7 offset:1024\n\t"
        "ds_read_b128 %4, %7 offset:2048\n\t"
        "ds_read_b128 %5, %7 offset:3072\n\t"
        "s_waitcnt lgkmcnt(2)"
        : "=&v"(a[0]), "=&v"(a[1]), "=&v"(b[0]), "=&v"(b[1]), "=&v"(b[2]), "=&v"(b[3])
        : "v"(bufb + offA), "v"(bufb + offB)
        : "memory");
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    asm volatile("s_waitcnt lgkmcnt(0)" : "+v"(b[2]), "+v"(b[3]) : : "memory");
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 2; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
  }
  __syncthreads();  // every wave is done with the ring before the epilogue reuses it
  if (nzb) {
    // compact rows: only the query's non-zero items, at their rank among them.  The stores scatter anyway,
    // so each lane writes its accumulators directly: row kg*4 + t of block i, column j*16 + r16 of the wave's
    // 64, which lie in bit words wn*2 (j < 2) and wn*2 + 1 (j >= 2) of the tile's four
    nz_bits[tid] = nzv[0];
    nz_rank[tid] = nzv[1];
    __syncthreads();
    const uint32_t wd0 = (uint32_t)(wn * 2);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const uint32_t tr = (uint32_t)(wm * 32 + i * 16 + kg * 4 + t);
        const uint64_t q = qbase + tr;
        if (q >= qn) continue;
        const uint32_t w2[2] = {nz_bits[tr * 4 + wd0], nz_bits[tr * 4 + wd0 + 1]};
        const uint32_t p2[2] = {nz_rank[tr * 4 + wd0], nz_rank[tr * 4 + wd0 + 1]};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const uint32_t word = w2[j >> 1], bit = (uint32_t)((j & 1) * 16 + r16);
          const uint64_t c = cbase + wn * 64 + j * 16 + r16;
          if (c < C && ((word >> bit) & 1u))
            S[q * ld + p2[j >> 1] + __popc(word & ((1u << bit) - 1u))] = cn[j] - 2.0f * acc[i][j][t];
        }
      }
    }
    return;
  }
  // dense rows: each wave passes its 32 x 64 block through LDS 16 rows at a time, so that every store
  // writes 256 contiguous bytes of a row of S
  float* ep = reinterpret_cast<float*>(ring) + wave * 16 * 65;
  const uint64_t cw = cbase + wn * 64 + lane;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int t = 0; t < 4; ++t) ep[(kg * 4 + t) * 65 + j * 16 + r16] = cn[j] - 2.0f * acc[i][j][t];
    __syncthreads();
    for (int r = 0; r < 16; ++r) {
      const uint64_t q = qbase + wm * 32 + i * 16 + r;
      if (q < qn && cw < C) S[q * ld + cw] = ep[r * 65 + lane];
    }
    __syncthreads();
  }
}

// |q|^2 of a d-dim query by one wave: lane partials over t = lane, lane + 64, ... in order, then a butterfly
// over the 64 lanes; every kernel that needs the value calls this, so they agree bit for bit.
// The empty asm takes the value as an operand, so the compiler keeps the load that produces it unconditional
// and ahead of the statement: a batch of loads followed by hold() on each is in flight together, where loads
// guarded by per-element conditions would each wait for the one before.
__device__ __forceinline__ void hold(float& x) { asm volatile("" : "+v"(x)); }
__device__ __forceinline__ void hold(uint32_t& x) { asm volatile("" : "+v"(x)); }

// wave_sq_norm in two halves for rows of at most 512 dims, so that a kernel can issue the row's loads with its
// other first loads and finish the sum later: lane partials over t = lane, lane + 64, ... in order, then a
// butterfly over the 64 lanes.
__device__ __forceinline__ void sq_norm_load(const float* q, uint64_t d, int lane, float (&x)[8]) {
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const uint64_t t = (uint64_t)lane + 64u * i;
    x[i] = q[t < d ? t : d - 1];
  }
}

__device__ __forceinline__ float sq_norm_finish(float (&x)[8], uint64_t d, int lane) {
#pragma unroll
  for (int i = 0; i < 8; ++i) hold(x[i]);
  float p = 0.0f;
#pragma unroll
  for (int i = 0; i < 8; ++i)
    if ((uint64_t)lane + 64u * i < d) p += x[i] * x[i];
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) p += __shfl_xor(p, m, 64);
  return __shfl(p, 0, 64);
}

__device__ __forceinline__ float wave_sq_norm(const float* q, uint64_t d, int lane) {
  if (d <= 512) {
    float x[8];
    sq_norm_load(q, d, lane, x);
    return sq_norm_finish(x, d, lane);
  }
  float p = 0.0f;
  for (uint64_t t0 = (uint64_t)lane; t0 < d; t0 += 512) {
    float x[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const uint64_t t = t0 + 64u * i;
      x[i] = q[t < d ? t : d - 1];
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) hold(x[i]);
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (t0 + 64u * i < d) p += x[i] * x[i];
  }
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) p += __shfl_xor(p, m, 64);
  return __shfl(p, 0, 64);
}

// score of listed item c in its query's score row: dense rows hold every item at its index, compact rows
// (nzb given) only the query's non-zero items, at their rank (word prefix nzp + bits below in the word)
__device__ __forceinline__ float item_score(const float* row, uint32_t c, const uint32_t* nzb_q,
                                            const uint32_t* nzp_q) {
  if (!nzb_q) return row[c];
  const uint32_t word = nzb_q[c >> 5], b = c & 31u;
  return row[nzp_q[c >> 5] + __popc(word & ((1u << b) - 1u))];
}

// order-preserving u32 key of a signed f32
__device__ __forceinline__ uint32_t ord(float v) {
  uint32_t u = __float_as_uint(v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Top-k of a query's listed items by their row of scores (ascending, ties to the lower list position),
// written as positions in the list: the mixed-precision replacement of the exact-FP32 refine + select_k
// used for cluster (A5) and sub-cluster (A6) selection.  k = budget[q] when given (0 = k_default), clamped to k_default.
// Optionally writes the 32 smallest listed scores (+ |q|^2) as budget-model features.
__global__ __launch_bounds__(256) void scored_select_kernel(const float* __restrict__ queries, uint64_t q0,
                                                            uint64_t d, const uint32_t* __restrict__ list,
                                                            uint32_t cap, const uint32_t* __restrict__ n_list_q,
                                                            const float* __restrict__ S, uint64_t ld,
                                                            uint32_t k_default, const uint32_t* __restrict__ budget,
                                                            uint32_t out_ld, uint32_t* __restrict__ chosen,
                                                            uint32_t* __restrict__ n_chosen,
                                                            float* __restrict__ feat_dist,
                                                            uint64_t* __restrict__ scratch, uint32_t scap,
                                                            int64_t skip_upto, const uint32_t* __restrict__ nzb,
                                                            const uint32_t* __restrict__ nzp, uint32_t nw,
                                                            const uint32_t* __restrict__ abort_flag, bool by_pos) {
  if (abort_flag && *abort_flag) return;
  extern __shared__ __align__(16) uint64_t keys[];  // scap keys
  __shared__ __align__(16) uint32_t hist[2048];
  __shared__ uint64_t sh[8];
  __shared__ float qq_sh;
  const uint64_t qi = blockIdx.x;
  const uint32_t n = n_list_q[qi];
  if (n <= skip_upto) return;  // handled by scored_select_wave_kernel
  const uint32_t* lst = list + qi * cap;
  const float* row = S + qi * ld;
  const uint32_t* nzb_q = nzb ? nzb + qi * nw : nullptr;
  const uint32_t* nzp_q = nzb ? nzp + qi * nw : nullptr;
  uint32_t k = k_default;
  if (budget) {
    const uint32_t b = budget[qi];
    if (b != 0 && b < k) k = b;
  }
  uint32_t m = n < k ? n : k;
  // keys selected: the m written, and at least the 32 smallest when the features are written
  uint32_t msel = m;
  if (feat_dist && msel < (n < 32u ? n : 32u)) msel = n < 32u ? n : 32u;
  if (n <= scap) {
    for (uint32_t j = threadIdx.x; j < n; j += kWG)
      keys[j] = ((uint64_t)ord(by_pos ? row[j] : item_score(row, lst[j], nzb_q, nzp_q)) << 32) | j;
    __syncthreads();
    dev::wg_sort_lds(keys, n);
  } else {
    // the keys and their range go to global scratch; one histogram pass over 2048 bins of the range and one
    // split pass (keys in bins below the msel-th key's bin to LDS, keys in its bin compacted in place at the
    // front of the scratch) leave a radix selection inside that one bin.  The msel smallest keys are unique,
    // so the result equals a full sort's first msel.
    __shared__ uint32_t mm[2][kWG / 64];
    uint64_t* g = scratch + qi * cap;
    uint32_t mn = 0xFFFFFFFFu, mx = 0u;
    for (uint32_t j = threadIdx.x; j < n; j += kWG) {
      const uint32_t o = ord(by_pos ? row[j] : item_score(row, lst[j], nzb_q, nzp_q));
      g[j] = ((uint64_t)o << 32) | j;
      mn = o < mn ? o : mn;
      mx = o > mx ? o : mx;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const uint32_t a = (uint32_t)__shfl_xor((int)mn, o, 64), b = (uint32_t)__shfl_xor((int)mx, o, 64);
      mn = a < mn ? a : mn;
      mx = b > mx ? b : mx;
    }
    if ((threadIdx.x & 63) == 0) {
      mm[0][threadIdx.x >> 6] = mn;
      mm[1][threadIdx.x >> 6] = mx;
    }
    for (int i = threadIdx.x; i < 2048; i += kWG) hist[i] = 0;
    __syncthreads();
    mn = mm[0][0];
    mx = mm[1][0];
#pragma unroll
    for (int w = 1; w < kWG / 64; ++w) {
      mn = mm[0][w] < mn ? mm[0][w] : mn;
      mx = mm[1][w] > mx ? mm[1][w] : mx;
    }
    if (msel > scap) msel = scap;
    if (m > msel) m = msel;
    const uint32_t sft = dev::dist_bin_shift(mn, mx);
    for (uint32_t j = threadIdx.x; j < n; j += kWG) atomicAdd(&hist[((uint32_t)(g[j] >> 32) - mn) >> sft], 1u);
    __syncthreads();
    dev::wg_hist_find(hist, msel, sh);  // sh[0] = the msel-th key's bin, sh[1] = keys below it
    uint32_t* cnt = reinterpret_cast<uint32_t*>(&sh[2]);
    if (threadIdx.x == 0) {
      cnt[0] = 0;
      cnt[1] = 0;
    }
    __syncthreads();
    const uint32_t bstar = (uint32_t)sh[0], below = (uint32_t)sh[1];
    for (uint32_t r0 = 0; r0 < n; r0 += kWG) {
      const uint32_t j = r0 + threadIdx.x;
      const uint64_t key = j < n ? g[j] : 0ull;
      // every key of this round is read before any is written back; writes land below r0 + kWG, reads of the
      // next rounds at or above it
      __syncthreads();
      if (j < n) {
        const uint32_t b = ((uint32_t)(key >> 32) - mn) >> sft;
        if (b < bstar) keys[atomicAdd(&cnt[0], 1u)] = key;
        else if (b == bstar) g[atomicAdd(&cnt[1], 1u)] = key;
      }
    }
    __syncthreads();
    const uint32_t mb = cnt[1], need = msel - below;
    __syncthreads();  // sh is reused by the radix selection
    if (need == mb) {
      for (uint32_t i = threadIdx.x; i < mb; i += kWG) keys[below + i] = g[i];
    } else {
      dev::wg_radix_select(dev::KeyArray{g}, mb, need, keys + below, hist, sh);
    }
    __syncthreads();
    dev::wg_sort_lds(keys, msel);
  }
  for (uint32_t j = threadIdx.x; j < m; j += kWG) chosen[qi * out_ld + j] = (uint32_t)keys[j];
  if (threadIdx.x == 0) n_chosen[qi] = m;
  if (feat_dist) {
    if (threadIdx.x < 64) {
      const float qq = wave_sq_norm(queries + (q0 + qi) * d, d, threadIdx.x);
      if (threadIdx.x == 0) qq_sh = qq;
    }
    __syncthreads();
    const uint32_t avail = n < 32 ? n : 32;
    for (uint32_t j = threadIdx.x; j < (uint32_t)kFeatDists; j += kWG) {
      float v = 0.0f;
      if (avail > 0) {
        const uint32_t o = (uint32_t)(keys[j < avail ? j : avail - 1] >> 32);
        const uint32_t u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
        v = __uint_as_float(u) + qq_sh;
      }
      feat_dist[qi * kFeatDists + j] = v;
    }
  }
}

// Wave-per-query variant of scored_select_kernel for lists of at most 64 * R items: the keys stay in
// registers, element e = lane * R + r (consecutive elements in one lane), and a bitonic network sorts them
// with register swaps for partners in the same lane (j < R) and shuffles for partners in other lanes.  No LDS,
// no barriers.  It sorts the same keys as the workgroup kernel, so the outputs are identical.  Lists longer
// than kWaveSelMax items are left to the workgroup kernel.
constexpr uint32_t kWaveSelMax = 512;

// Sorts the c (<= 64 * R2) keys in lbuf and writes the first m indexes; feat_key as in wave_select.
template <int R2>
__device__ __forceinline__ void wave_sort_out(const uint64_t* lbuf, uint32_t c, uint32_t m, int lane, uint32_t* out,
                                              uint32_t avail, uint64_t& feat_key) {
  uint64_t key[R2];
#pragma unroll
  for (int r = 0; r < R2; ++r) {
    const uint32_t e = (uint32_t)lane * R2 + r;
    key[r] = e < c ? lbuf[e] : ~0ull;
  }
  dev::wave_bitonic<R2>(key, lane);
#pragma unroll
  for (int r = 0; r < R2; ++r) {
    const uint32_t e = (uint32_t)lane * R2 + r;
    if (e < m) out[e] = (uint32_t)key[r];
  }
  if (avail > 0) {
    const uint32_t j = (uint32_t)lane < avail ? (uint32_t)lane : avail - 1;
    feat_key = dev::wave_key_at<R2>(key, j & 63u);
  }
}

// The m_need smallest of a wave's keys (key = ord(score) << 32 | element, ~0 = pad; keys are distinct) without
// sorting all of them: a radix descent over the 32-bit score order, one byte a step (a 256-bin histogram of the
// keys inside the current range in the wave's LDS, a scan of it across the lanes), narrows the score range
// holding the m_need-th smallest key until the keys below the end of that range number at most `target`; those
// keys (a superset of the m_need smallest: every key below the range is smaller, every key above it larger) are
// compacted into lbuf in element order.  Returns their count, or 0 when more than `target` keys tie at one
// score (the caller then sorts everything).
template <int R>
__device__ __forceinline__ uint32_t wave_cut_smallest(const uint64_t (&key)[R], uint32_t n, uint32_t m_need,
                                                      uint32_t target, int lane, uint64_t* lbuf, uint32_t* hist) {
  uint32_t below = 0, inb = n;  // valid keys below the range, in the range
  uint64_t prefix = 0;          // range = the keys whose score order >> shift equals prefix
  int shift = 32;
  if (n > target) {
    // the bits above the highest one in which the scores differ are common: start below them, so that the
    // first histogram spreads over its bins instead of piling every key into one
    uint32_t mn = 0xFFFFFFFFu, mx = 0u;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (key[r] != ~0ull) {
        const uint32_t o = (uint32_t)(key[r] >> 32);
        mn = o < mn ? o : mn;
        mx = o > mx ? o : mx;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const uint32_t a = (uint32_t)__shfl_xor((int)mn, o, 64), b = (uint32_t)__shfl_xor((int)mx, o, 64);
      mn = a < mn ? a : mn;
      mx = b > mx ? b : mx;
    }
    shift = 32 - __clz((int)(mn ^ mx));  // 0 when every score is equal
    prefix = (uint64_t)mx >> shift;
  }
  while (below + inb > target) {
    if (shift == 0) return 0;
    const int wd = shift < 8 ? shift : 8;
    shift -= wd;
    reinterpret_cast<uint4*>(hist)[lane] = make_uint4(0u, 0u, 0u, 0u);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint64_t o = key[r] >> 32;
      if (key[r] != ~0ull && (o >> (shift + wd)) == prefix)
        atomicAdd(&hist[(uint32_t)(o >> shift) & ((1u << wd) - 1u)], 1u);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const uint4 h = reinterpret_cast<const uint4*>(hist)[lane];  // bins 4 lane .. 4 lane + 3
    const uint32_t s0 = h.x, s1 = s0 + h.y, s2 = s1 + h.z, s3 = s2 + h.w;
    uint32_t incl = s3;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)incl, o, 64);
      if (lane >= o) incl += y;
    }
    const uint32_t ex = incl - s3;
    const uint32_t need = m_need - below;  // rank (from 1) of the m_need-th smallest inside the range
    // the one lane whose bins take the cumulative count from below `need` to at least `need`
    const bool mine = ex < need && incl >= need;
    const uint64_t bal = __ballot(mine);
    const int src = __ffsll((unsigned long long)bal) - 1;
    uint32_t lo = ex, hi = ex + s0, dg = 0;
    if (ex + s0 < need) {
      lo = ex + s0; hi = ex + s1; dg = 1;
      if (ex + s1 < need) {
        lo = ex + s1; hi = ex + s2; dg = 2;
        if (ex + s2 < need) {
          lo = ex + s2; hi = ex + s3; dg = 3;
        }
      }
    }
    const uint32_t bin = (uint32_t)__shfl((int)((uint32_t)lane * 4u + dg), src, 64);
    const uint32_t blo = (uint32_t)__shfl((int)lo, src, 64), bhi = (uint32_t)__shfl((int)hi, src, 64);
    below += blo;
    inb = bhi - blo;
    prefix = (prefix << wd) | bin;
    __builtin_amdgcn_wave_barrier();  // every lane has read the histogram before the next step clears it
  }
  // compaction of the keys with score order < (prefix + 1) << shift, element order kept
  bool keep[R];
  uint32_t cnt = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    keep[r] = key[r] != ~0ull && ((key[r] >> 32) >> shift) <= prefix;
    cnt += keep[r] ? 1u : 0u;
  }
  uint32_t incl = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)incl, o, 64);
    if (lane >= o) incl += y;
  }
  uint32_t off = incl - cnt;
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (keep[r]) lbuf[off++] = key[r];
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  return below + inb;
}

template <int R>
__device__ void wave_select(const uint32_t* lst, const float* row, const uint32_t* nzb_q, const uint32_t* nzp_q,
                            bool by_pos, uint32_t n, uint32_t m, int lane, uint32_t* out, uint32_t avail,
                            uint64_t& feat_key, uint64_t* lbuf, uint32_t* hist) {
  // every load unconditional (elements past n read element n - 1), so the wave has all of them in flight
  // at once instead of one guarded load chain per element
  uint64_t key[R];
#pragma unroll
  for (int r = 0; r < R; ++r) key[r] = ~0ull;
  if (n > 0) {
    float v[R];
    if (by_pos) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const uint32_t e = (uint32_t)lane * R + r;
        v[r] = row[e < n ? e : n - 1];
      }
    } else {
      // item_score's loads in three batches: list items, their bit words and word prefixes, the scores
      uint32_t id[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const uint32_t e = (uint32_t)lane * R + r;
        id[r] = lst[e < n ? e : n - 1];
      }
#pragma unroll
      for (int r = 0; r < R; ++r) hold(id[r]);
      if (nzb_q) {
        uint32_t wd[R], pf[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          wd[r] = nzb_q[id[r] >> 5];
          pf[r] = nzp_q[id[r] >> 5];
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
          hold(wd[r]);
          hold(pf[r]);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = row[pf[r] + __popc(wd[r] & ((1u << (id[r] & 31u)) - 1u))];
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = row[id[r]];
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) hold(v[r]);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t e = (uint32_t)lane * R + r;
      if (e < n) key[r] = ((uint64_t)ord(v[r]) << 32) | e;
    }
  }
  if constexpr (R >= 4) {
    // fewer wanted than held: cut to a smaller sort
    const uint32_t m_need = m > avail ? m : avail;
    const uint32_t target = m_need <= 64u ? 64u : (m_need <= 128u ? 128u : 256u);
    if (m_need > 0 && target < 64u * R) {
      const uint32_t c = wave_cut_smallest<R>(key, n, m_need, target, lane, lbuf, hist);
      if (c > 0) {
        if (c <= 64u) wave_sort_out<1>(lbuf, c, m, lane, out, avail, feat_key);
        else if (c <= 128u) wave_sort_out<2>(lbuf, c, m, lane, out, avail, feat_key);
        else wave_sort_out<4>(lbuf, c, m, lane, out, avail, feat_key);
        return;
      }
    }
  }
  dev::wave_bitonic<R>(key, lane);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)lane * R + r;
    if (e < m) out[e] = (uint32_t)key[r];
  }
  if (avail > 0) {
    // lane j < 32 takes sorted key min(j, avail - 1)
    const uint32_t j = (uint32_t)lane < avail ? (uint32_t)lane : avail - 1;
    feat_key = dev::wave_key_at<R>(key, j & 63u);
  }
}

__global__ __launch_bounds__(256) void scored_select_wave_kernel(
    const float* __restrict__ queries, uint64_t q0, uint64_t qn, uint64_t d, const uint32_t* __restrict__ list,
    uint32_t cap, const uint32_t* __restrict__ n_list_q, const float* __restrict__ S, uint64_t ld,
    uint32_t k_default, const uint32_t* __restrict__ budget, uint32_t out_ld, uint32_t* __restrict__ chosen,
    uint32_t* __restrict__ n_chosen, float* __restrict__ feat_dist, const uint32_t* __restrict__ nzb,
    const uint32_t* __restrict__ nzp, uint32_t nw, uint32_t* __restrict__ abort_flag, uint32_t* __restrict__ left,
    bool abort_on_left, bool by_pos) {
  if (abort_flag && *abort_flag) return;
  const int lane = threadIdx.x & 63;
  const uint64_t qi = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (qi >= qn) return;
  const uint32_t n = n_list_q[qi];
  if (n > kWaveSelMax) {
    // left to the workgroup kernel: counted; a replayed graph that left that kernel out is abandoned
    if (lane == 0) {
      if (left) atomicAdd(left, 1u);
      if (abort_on_left) *abort_flag = 1u;
    }
    return;
  }
  const uint32_t* lst = list + qi * cap;
  const float* row = S + qi * ld;
  const uint32_t* nzb_q = nzb ? nzb + qi * nw : nullptr;
  const uint32_t* nzp_q = nzb ? nzp + qi * nw : nullptr;
  uint32_t k = k_default;
  if (budget) {
    const uint32_t b = budget[qi];
    if (b != 0 && b < k) k = b;
  }
  const uint32_t m = n < k ? n : k;
  uint32_t* out = chosen + qi * out_ld;
  const uint32_t avail = feat_dist ? (n < 32 ? n : 32) : 0;
  uint64_t fk = 0;
  __shared__ uint64_t lbuf_all[4][256];
  __shared__ __align__(16) uint32_t hist_all[4][256];
  uint64_t* lbuf = lbuf_all[threadIdx.x >> 6];
  uint32_t* hist = hist_all[threadIdx.x >> 6];
  if (n <= 64) wave_select<1>(lst, row, nzb_q, nzp_q, by_pos, n, m, lane, out, avail, fk, lbuf, hist);
  else if (n <= 128) wave_select<2>(lst, row, nzb_q, nzp_q, by_pos, n, m, lane, out, avail, fk, lbuf, hist);
  else if (n <= 256) wave_select<4>(lst, row, nzb_q, nzp_q, by_pos, n, m, lane, out, avail, fk, lbuf, hist);
  else wave_select<8>(lst, row, nzb_q, nzp_q, by_pos, n, m, lane, out, avail, fk, lbuf, hist);
  if (lane == 0) n_chosen[qi] = m;
  if (feat_dist) {
    const float qq = wave_sq_norm(queries + (q0 + qi) * d, d, lane);
    if (lane < kFeatDists) {
      float v = 0.0f;
      if (avail > 0) {
        const uint32_t o = (uint32_t)(fk >> 32);
        const uint32_t u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
        v = __uint_as_float(u) + qq;
      }
      feat_dist[qi * kFeatDists + lane] = v;
    }
  }
}

// The c (<= 64 * R) keys in buf sorted; returns the lane-th smallest (~0 past c).
template <int R>
__device__ __forceinline__ uint32_t wave_sorted_at_lane(const uint32_t* buf, uint32_t c, int lane) {
  uint32_t key[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)lane * R + r;
    key[r] = e < c ? buf[e] : 0xFFFFFFFFu;
  }
  dev::wave_bitonic32<R>(key, lane);
  // rank j lies in lane j / R, register j % R
  const uint32_t src = (uint32_t)lane / R, rr = (uint32_t)lane % R;
  uint32_t v = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t y = (uint32_t)__shfl((int)key[r], (int)src, 64);
    v = (uint32_t)r == rr ? y : v;
  }
  return v;
}

// Budget-model features on the path that probes every non-zero cluster: per query, the 32 smallest coarse
// scores among its non-zero clusters, ascending, plus |q|^2 (lane j < 32 writes the min(j, avail - 1)-th,
// avail = min(n, 32)) -- the values scored_select_wave_kernel writes as its features, without its full sort of
// 64-bit (score, position) keys.  One wave per query, lists of at most 512 items, element e = r * 64 + lane (the
// list and score loads coalesced, every lane holding an item once n >= 64): T = the 32nd smallest of the 64
// lane minima, so at least 32 keys are <= T and the 32 smallest are among them; those keys (typically ~n / 10)
// are compacted into LDS and sorted.  Keys are order-preserving u32 of the f32 scores, so equal scores give
// equal features whatever their order.
__global__ __launch_bounds__(256) void coarse_features_kernel(const float* __restrict__ queries, uint64_t q0,
                                                              uint64_t qn, uint64_t d,
                                                              const uint32_t* __restrict__ list, uint32_t cap,
                                                              const uint32_t* __restrict__ n_list_q,
                                                              const float* __restrict__ S, uint64_t ld,
                                                              float* __restrict__ feat,
                                                              const uint32_t* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  __shared__ uint32_t cbuf_all[4][512];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t qi = (uint64_t)blockIdx.x * 4 + wv;
  if (qi >= qn) return;
  uint32_t* cbuf = cbuf_all[wv];
  // every first load issued before any wait: the query row (for |q|^2), the list length and the list row
  // (read up to cap, as the length is not known yet), then the scores of the listed items
  const float* qrow = queries + (q0 + qi) * d;
  float qx[8];
  if (d <= 512) sq_norm_load(qrow, d, lane, qx);
  const uint32_t nraw = n_list_q[qi];
  const uint32_t* lst = list + qi * cap;
  uint32_t key[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const uint32_t e = (uint32_t)(r * 64 + lane);
    key[r] = lst[e < cap ? e : cap - 1];  // cap <= 512
  }
#pragma unroll
  for (int r = 0; r < 8; ++r) hold(key[r]);
  const uint32_t n = nraw < cap ? nraw : cap;
  const float* row = S + qi * ld;
  float x[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) x[r] = row[(uint32_t)(r * 64 + lane) < n ? key[r] : 0u];
#pragma unroll
  for (int r = 0; r < 8; ++r) hold(x[r]);
#pragma unroll
  for (int r = 0; r < 8; ++r) key[r] = (uint32_t)(r * 64 + lane) < n ? ord(x[r]) : 0xFFFFFFFFu;
  uint32_t v;  // lane j: the j-th smallest key (j < min(n, 64))
  if (n <= 64) {
    uint32_t k1[1] = {key[0]};
    dev::wave_bitonic32<1>(k1, lane);
    v = k1[0];
  } else {
    uint32_t mn = key[0];
#pragma unroll
    for (int r = 1; r < 8; ++r) mn = key[r] < mn ? key[r] : mn;
    uint32_t t[1] = {mn};
    dev::wave_bitonic32<1>(t, lane);
    const uint32_t T = (uint32_t)__shfl((int)t[0], 31, 64);
    uint32_t c = 0;
#pragma unroll
    for (int r = 0; r < 8; ++r) c += key[r] <= T ? 1u : 0u;
    uint32_t incl = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)incl, o, 64);
      if (lane >= o) incl += y;
    }
    const uint32_t total = (uint32_t)__shfl((int)incl, 63, 64);
    uint32_t off = incl - c;
#pragma unroll
    for (int r = 0; r < 8; ++r)
      if (key[r] <= T) cbuf[off++] = key[r];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (total <= 64) v = wave_sorted_at_lane<1>(cbuf, total, lane);
    else if (total <= 128) v = wave_sorted_at_lane<2>(cbuf, total, lane);
    else if (total <= 256) v = wave_sorted_at_lane<4>(cbuf, total, lane);
    else v = wave_sorted_at_lane<8>(cbuf, total, lane);
  }
  const uint32_t avail = n < (uint32_t)kFeatDists ? n : (uint32_t)kFeatDists;
  const uint32_t j = avail == 0 ? 0u : ((uint32_t)lane < avail ? (uint32_t)lane : avail - 1);
  const uint32_t o = (uint32_t)__shfl((int)v, (int)j, 64);
  const float qq = d <= 512 ? sq_norm_finish(qx, d, lane) : wave_sq_norm(qrow, d, lane);
  if (lane < kFeatDists) {
    const uint32_t u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
    feat[qi * kFeatDists + lane] = avail > 0 ? __uint_as_float(u) + qq : 0.0f;
  }
}

}  // namespace

void launch_scored_select(const float* queries, uint64_t q0, uint64_t qn, uint64_t d, const uint32_t* list,
                          uint32_t cap, const uint32_t* n_list_q, const float* S, uint64_t ld, uint32_t k_default,
                          const uint32_t* budget, uint32_t out_ld, uint32_t* chosen, uint32_t* n_chosen,
                          float* feat_dist, uint64_t* scratch, bool wave, hipStream_t s, const uint32_t* nzb,
                          const uint32_t* nzp, uint32_t nw, uint32_t* abort_flag, uint32_t* left,
                          bool skip_fallback, bool by_pos) {
  uint32_t scap = 1024;
  while (scap < k_default) scap <<= 1;
  int64_t skip_upto = -1;
  if (wave) {
    hipLaunchKernelGGL(scored_select_wave_kernel, dim3((unsigned)((qn + 3) / 4)), dim3(256), 0, s, queries, q0, qn,
                       d, list, cap, n_list_q, S, ld, k_default, budget, out_ld, chosen, n_chosen, feat_dist, nzb,
                       nzp, nw, abort_flag, left, skip_fallback && abort_flag, by_pos);
    skip_upto = kWaveSelMax;
    if (skip_fallback && abort_flag) return;  // captured without the workgroup kernel
  }
  hipLaunchKernelGGL(scored_select_kernel, dim3((unsigned)qn), dim3(256), scap * 8, s, queries, q0, d, list, cap,
                     n_list_q, S, ld, k_default, budget, out_ld, chosen, n_chosen, feat_dist, scratch, scap,
                     skip_upto, nzb, nzp, nw, abort_flag, by_pos);
}

void launch_coarse_features(const float* queries, uint64_t q0, uint64_t qn, uint64_t d, const uint32_t* list,
                            uint32_t cap, const uint32_t* n_list_q, const float* S, uint64_t ld, float* feat,
                            hipStream_t s, const uint32_t* abort_flag) {
  if (cap > 512) fail(kUnsupported, "coarse features take lists of at most 512 items");
  if (qn == 0) return;
  hipLaunchKernelGGL(coarse_features_kernel, dim3((unsigned)((qn + 3) / 4)), dim3(256), 0, s, queries, q0, qn, d,
                     list, cap, n_list_q, S, ld, feat, abort_flag);
}

// 8 elements per thread (two float4 loads, one 16-byte store) when d and Dp are multiples of 8
__global__ void to_bf16_x8_kernel(const float* __restrict__ queries, uint64_t q0, uint64_t qn, uint64_t d,
                                  uint64_t Dp, uint16_t* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;  // 8-element group
  const uint64_t per = Dp / 8;
  if (i >= qn * per) return;
  const uint64_t q = i / per, t = (i % per) * 8;
  uint32_t pk[4] = {0u, 0u, 0u, 0u};
  if (t < d) {
    const float4 a = *reinterpret_cast<const float4*>(queries + (q0 + q) * d + t);
    const float4 b = *reinterpret_cast<const float4*>(queries + (q0 + q) * d + t + 4);
    const float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int k = 0; k < 4; ++k)
      pk[k] = (uint32_t)(uint16_t)f32_to_bf16(v[2 * k]) | ((uint32_t)(uint16_t)f32_to_bf16(v[2 * k + 1]) << 16);
  }
  *reinterpret_cast<uint4*>(out + q * Dp + t) = make_uint4(pk[0], pk[1], pk[2], pk[3]);
}

void launch_to_bf16(const float* queries, uint64_t q0, uint64_t qn, uint64_t d, uint64_t Dp, uint16_t* out,
                    hipStream_t s) {
  if (d % 8 == 0 && Dp % 8 == 0) {
    const uint64_t n = qn * (Dp / 8);
    hipLaunchKernelGGL(to_bf16_x8_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, queries, q0, qn, d, Dp,
                       out);
    return;
  }
  const uint64_t n = qn * Dp;
  hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, queries, q0, qn, d, Dp, out);
}

void launch_coarse_gemm(const uint16_t* qb16, uint64_t qn, uint64_t Dp, const uint16_t* cb16, const float* cnorm,
                        uint32_t C, float* S, hipStream_t s, const uint32_t* nzb, const uint32_t* nzp, uint32_t nw,
                        const uint32_t* abort_flag) {
  const uint64_t QT = (qn + kGemmTileQ - 1) / kGemmTileQ, CT = (C + kGemmTile - 1) / kGemmTile;
  if (QT * CT > 0xFFFFFFFFull / 2) fail(kUnsupported, "coarse GEMM grid too large");
  const uint32_t per = (uint32_t)((QT * CT + kXcds - 1) / kXcds);
  hipLaunchKernelGGL(coarse_gemm_kernel, dim3(per * kXcds), dim3(kGemmThreads), 0, s, qb16, qn, Dp, cb16, cnorm, C, S,
                     (uint32_t)QT, (uint32_t)CT, per, nzb, nzp, nw, (uint64_t)C, abort_flag);
}

}  // namespace hdg
