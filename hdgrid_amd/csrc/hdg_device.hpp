// Device building blocks shared by the query kernels: 256-thread workgroup scans, LDS bitonic sort,
// radix selection over unique 64-bit keys, and the exact FP32 L2 in the reference's order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hdg {
namespace dev {

constexpr int kWG = 256;              // threads per workgroup for the per-query kernels
constexpr int kSortCap = 4096;        // keys sorted in LDS (32 KiB)
constexpr uint64_t kKeyPad = ~0ull;

// Unique, totally ordered key for (non-negative f32 distance, 32-bit tie index).  For x >= 0 (and
// +inf / NaN, which sort last) the IEEE bit pattern is monotone.
__device__ __forceinline__ uint64_t make_key(float v, uint32_t tie) {
  return ((uint64_t)__float_as_uint(v) << 32) | tie;
}
__device__ __forceinline__ float key_value(uint64_t k) { return __uint_as_float((uint32_t)(k >> 32)); }
__device__ __forceinline__ uint32_t key_tie(uint64_t k) { return (uint32_t)k; }

// Exclusive prefix sum over the 256 threads of a workgroup; `total` receives the sum.
// `tmp` must hold kWG/64 + 1 words of LDS.
__device__ __forceinline__ uint32_t wg_exclusive_scan(uint32_t v, uint32_t* tmp, uint32_t& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) tmp[wave] = x;
  __syncthreads();
  uint32_t base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kWG / 64; ++w) {
    uint32_t t = tmp[w];
    if (w < wave) base += t;
    tot += t;
  }
  __syncthreads();
  total = tot;
  return base + x - v;
}

// Bitonic sort of keys[0..n) in LDS ascending (n <= kSortCap); pads with kKeyPad up to a power of 2.
__device__ inline void wg_sort_lds(uint64_t* keys, uint32_t n) {
  uint32_t np = 1;
  while (np < n) np <<= 1;
  for (uint32_t i = n + threadIdx.x; i < np; i += kWG) keys[i] = kKeyPad;
  __syncthreads();
  for (uint32_t k = 2; k <= np; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t t = threadIdx.x; t < np / 2; t += kWG) {
        uint32_t i = 2 * t - (t & (j - 1));
        uint32_t p = i + j;
        bool up = (i & k) == 0;
        uint64_t a = keys[i], b = keys[p];
        if ((a > b) == up) {
          keys[i] = b;
          keys[p] = a;
        }
      }
      __syncthreads();
    }
  }
}

// 64-bit shuffles (two 32-bit halves)
__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src) {
  const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, 64);
  const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, 64);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
  const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m, 64);
  const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m, 64);
  return ((uint64_t)hi << 32) | lo;
}

// Bitonic sort of 64 * R keys held by one wave in registers, element e = lane * R + r (consecutive elements
// in one lane): register swaps for partners in the same lane (j < R), shuffles for partners in other lanes.
// Ascending; no LDS, no barriers.
template <int R>
__device__ __forceinline__ void wave_bitonic(uint64_t (&key)[R], int lane) {
  constexpr uint32_t N = 64u * R;
#pragma unroll
  for (uint32_t k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      if (j < (uint32_t)R) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int p = r ^ (int)j;
          if (p > r) {
            const uint32_t e = (uint32_t)lane * R + r;
            const bool up = (e & k) == 0;
            const uint64_t a = key[r], b = key[p];
            if ((a > b) == up) {
              key[r] = b;
              key[p] = a;
            }
          }
        }
      } else {
        const int lm = (int)(j / R);
        const bool low = (lane & lm) == 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const uint32_t e = (uint32_t)lane * R + r;
          const uint64_t o = shfl_xor_u64(key[r], lm);
          const bool up = (e & k) == 0;
          // the lower element of the pair keeps the smaller key when ascending
          const uint64_t mn = key[r] < o ? key[r] : o, mx = key[r] < o ? o : key[r];
          key[r] = (low == up) ? mn : mx;
        }
      }
    }
  }
}

// The same network for 32-bit keys.
template <int R>
__device__ __forceinline__ void wave_bitonic32(uint32_t (&key)[R], int lane) {
  constexpr uint32_t N = 64u * R;
#pragma unroll
  for (uint32_t k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      if (j < (uint32_t)R) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int p = r ^ (int)j;
          if (p > r) {
            const uint32_t e = (uint32_t)lane * R + r;
            const bool up = (e & k) == 0;
            const uint32_t a = key[r], b = key[p];
            if ((a > b) == up) {
              key[r] = b;
              key[p] = a;
            }
          }
        }
      } else {
        const int lm = (int)(j / R);
        const bool low = (lane & lm) == 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const uint32_t e = (uint32_t)lane * R + r;
          const uint32_t o = (uint32_t)__shfl_xor((int)key[r], lm, 64);
          const bool up = (e & k) == 0;
          const uint32_t mn = key[r] < o ? key[r] : o, mx = key[r] < o ? o : key[r];
          key[r] = (low == up) ? mn : mx;
        }
      }
    }
  }
}

// sorted key number j (< 64 * R) of a wave-sorted array, broadcast from its lane
template <int R>
__device__ __forceinline__ uint64_t wave_key_at(const uint64_t (&key)[R], uint32_t j) {
  uint64_t v = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint64_t x = shfl_u64(key[r], (int)(j / R));
    if ((int)(j % R) == r) v = x;
  }
  return v;
}

// Radix selection: writes the k smallest of the unique keys g[0..n) (global memory) to out[0..k),
// unordered.  Requires k < n.  `hist` is 2048 words of LDS, `sh` 8 words of LDS.
// Key sources for the selections: a plain key array, or (f32 distance bits, index) built from a
// distance array.
struct KeyArray {
  const uint64_t* g;
  __device__ __forceinline__ uint64_t operator()(uint32_t i) const { return g[i]; }
};
struct KeyDist {
  const uint32_t* d;
  __device__ __forceinline__ uint64_t operator()(uint32_t i) const { return ((uint64_t)d[i] << 32) | i; }
};

// Bin holding the need-th smallest item (1-based, need <= total) of a 2048-bin LDS histogram, by an
// exclusive scan over 8 bins per thread: sh[0] = the bin, sh[1] = items in lower bins.  Ends with a barrier.
__device__ inline void wg_hist_find(const uint32_t* hist, uint32_t need, uint64_t* sh) {
  __shared__ uint32_t tmp[kWG / 64 + 1];
  static_assert(kWG * 8 == 2048, "8 bins per thread");
  const uint4 a = reinterpret_cast<const uint4*>(hist)[threadIdx.x * 2];
  const uint4 b = reinterpret_cast<const uint4*>(hist)[threadIdx.x * 2 + 1];
  const uint32_t c8[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  uint32_t s8 = 0;
#pragma unroll
  for (int t = 0; t < 8; ++t) s8 += c8[t];
  uint32_t total;
  uint32_t before = wg_exclusive_scan(s8, tmp, total);
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    if (before < need && before + c8[t] >= need) {
      sh[0] = threadIdx.x * 8 + t;
      sh[1] = before;
    }
    before += c8[t];
  }
  __syncthreads();
}

template <class Src>
__device__ inline void wg_radix_select(Src g, uint32_t n, uint32_t k, uint64_t* __restrict__ out, uint32_t* hist,
                                       uint64_t* sh) {
  uint64_t prefix = 0, mask = 0;
  uint32_t need = k;  // rank (1-based) of the k-th key inside the current prefix bucket
  const int shifts[6] = {53, 42, 31, 20, 9, 0};
  const int widths[6] = {11, 11, 11, 11, 11, 9};
#pragma unroll 1
  for (int pass = 0; pass < 6; ++pass) {
    const int sft = shifts[pass];
    const uint64_t dmask = (1ull << widths[pass]) - 1;
    for (int i = threadIdx.x; i < 2048; i += kWG) hist[i] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += kWG) {
      uint64_t key = g(i);
      if ((key & mask) == prefix) atomicAdd(&hist[(key >> sft) & dmask], 1u);
    }
    __syncthreads();
    wg_hist_find(hist, need, sh);
    const uint64_t b = sh[0];
    need -= (uint32_t)sh[1];
    prefix |= b << sft;
    mask |= dmask << sft;
    __syncthreads();
  }
  // prefix is now the k-th smallest key; keys are unique, so exactly k keys are <= prefix
  if (threadIdx.x == 0) sh[2] = 0;
  __syncthreads();
  uint32_t* slot = reinterpret_cast<uint32_t*>(&sh[2]);
  for (uint32_t i = threadIdx.x; i < n; i += kWG) {
    uint64_t key = g(i);
    if (key <= prefix) {
      uint32_t s = atomicAdd(slot, 1u);
      if (s < k) out[s] = key;  // exactly k qualify for unique keys; the guard bounds the write anyway
    }
  }
  __syncthreads();
}

// Bin shift for distance bits in [mn, mx]: (bits - mn) >> shift stays below 2048 and is monotone in the
// distance.
__device__ __forceinline__ uint32_t dist_bin_shift(uint32_t mn, uint32_t mx) {
  const uint32_t r = mx - mn;
  const int bits = r ? 32 - __clz((int)r) : 0;
  return bits > 11 ? (uint32_t)(bits - 11) : 0u;
}

// Distance items per thread kept in flight by the selection passes.
constexpr int kSelUnroll = 8;

// Selection of the k smallest keys (d[i] << 32 | i) of a distance array d[0..n) (non-negative f32 bits, all
// in [mn, mx]) by binning: one histogram pass into the 2048 bins of dist_bin_shift, one split pass (keys in
// bins below the k-th key's bin go to out[], keys in that bin to scratch[]), then a radix selection inside
// that one bin.  Bins are monotone in the distance, so out[0..k) holds exactly the k smallest keys
// (unordered), ties by index.  Requires k < n; hist = 2048 words of LDS, sh = 8 words of LDS, scratch = n
// keys.  U distance loads per thread stay in flight.
template <int U>
__device__ inline void wg_dist_select(const uint32_t* __restrict__ d, uint32_t n, uint32_t k, uint32_t mn,
                                      uint32_t mx, uint64_t* __restrict__ out, uint64_t* __restrict__ scratch,
                                      uint32_t* hist, uint64_t* sh) {
  constexpr uint32_t kStep = kWG * U;
  const uint32_t sft = dist_bin_shift(mn, mx);
  for (int i = threadIdx.x; i < 2048; i += kWG) hist[i] = 0;
  __syncthreads();
  for (uint32_t i0 = threadIdx.x; i0 < n; i0 += kStep) {
    uint32_t v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t i = i0 + (uint32_t)u * kWG;
      v[u] = i < n ? d[i] : 0u;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i0 + (uint32_t)u * kWG < n) atomicAdd(&hist[(v[u] - mn) >> sft], 1u);
  }
  __syncthreads();
  wg_hist_find(hist, k, sh);  // sh[0] = the k-th key's bin, sh[1] = keys below it
  uint32_t* cnt = reinterpret_cast<uint32_t*>(&sh[2]);  // [0] keys below, [1] keys in the bin
  if (threadIdx.x == 0) {
    cnt[0] = 0;
    cnt[1] = 0;
  }
  __syncthreads();
  const uint32_t bstar = (uint32_t)sh[0], below = (uint32_t)sh[1];
  for (uint32_t i0 = threadIdx.x; i0 < n; i0 += kStep) {
    uint32_t v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t i = i0 + (uint32_t)u * kWG;
      v[u] = i < n ? d[i] : 0u;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t i = i0 + (uint32_t)u * kWG;
      if (i < n) {
        const uint32_t b = (v[u] - mn) >> sft;
        const uint64_t key = ((uint64_t)v[u] << 32) | i;
        if (b < bstar) out[atomicAdd(&cnt[0], 1u)] = key;
        else if (b == bstar) scratch[atomicAdd(&cnt[1], 1u)] = key;
      }
    }
  }
  __syncthreads();
  const uint32_t m = cnt[1], need = k - below;
  __syncthreads();
  if (need == m) {
    for (uint32_t i = threadIdx.x; i < m; i += kWG) out[below + i] = scratch[i];
  } else {
    wg_radix_select(KeyArray{scratch}, m, need, out + below, hist, sh);
  }
  __syncthreads();
}

}  // namespace dev
}  // namespace hdg
