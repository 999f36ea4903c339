// Super-Grid query path on the GPU: Super-Grid-IVF_PQ.cu:985-1248 (index::query) and the kernels it
// reaches, redesigned for gfx950.  Stage map (reference row -> kernel here):
//   A1 filter recovery + A2 grid boxes + A7 Morton corners  -> prep_kernel
//   A3/A4 in-box counts, non-zero sub-clusters and clusters -> sparse_count_kernel (bucket-grid walk);
//                                                              dense_count_kernel + dense_emit_kernel for
//                                                              the boxes too wide for the walk
//   A5 exact coarse top-n_list                              -> pair_dist_kernel + coarse_pick_kernel
//   A6 non-zero sub-clusters of the chosen clusters, exact
//      top-sub_lists                                        -> sub_expand_kernel, pair_dist_kernel + sub_pick_kernel
//   A8/A9 Morton ranges, candidate counts, beta sums        -> ranges_kernel
//   A10 beta / n_candies, buffer plan                       -> plan_kernel + scan
//   A11 PQ ADC scan + top-k' (only where C_q > k')          -> adc_scan_kernel, adc_select_kernel
//   A12 label post-filter (l > 1)                           -> inside keep_all / adc_select
//   A13 exact FP32 re-rank + final top-k                    -> pair_dist_kernel + final_pick_kernel
// Arithmetic that decides integer results (cells, Morton codes, counts) and the FP32 distances follow
// the reference's operation order; the file is compiled with -ffp-contract=off and every FMA that the
// reference's compiler contracts is an explicit fmaf().
#include <hip/hip_runtime.h>

#include <mutex>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
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
constexpr int kCountSegChunk = 64;              // segments per dense_count_kernel block

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

// Zeroes the chunk's reference-batch statistics and flags (a kernel, so that a captured graph holds no
// memset node).
__global__ void clear_kernel(unsigned long long* __restrict__ gstat, uint64_t n, uint32_t* __restrict__ err,
                             uint32_t* __restrict__ n_dense, uint32_t* __restrict__ abort_flag) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) gstat[i] = 0;
  if (i == 0) {
    *err = 0;
    n_dense[0] = 0;  // queries handed to the dense count
    n_dense[1] = 0;  // queries handed from the wave walk to the workgroup walk
    n_dense[2] = 0;  // selection slots handed out by the plan
    n_dense[3] = 0;  // queries left to the workgroup scored selection
    n_dense[4] = 0;  // queries left to the workgroup final pick
    if (abort_flag) *abort_flag = 0;
  }
}

// Every query of the chunk goes to the dense count (no bucket grid, or too many sub-clusters for LDS).
__global__ void iota_kernel(uint64_t qn, uint32_t* __restrict__ dense_list, uint32_t* __restrict__ n_dense,
                            uint32_t* __restrict__ n_ib) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < qn) {
    dense_list[i] = (uint32_t)i;
    n_ib[i] = 0xFFFFFFFFu;
  }
  if (i == 0) *n_dense = (uint32_t)qn;
}

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
  const uint64_t za = morton_code(lo_c, l, D.max_bits, D.valid_map);
  zlo[qi] = za;
  zhi[qi] = morton_code(hi_c, l, D.max_bits, D.valid_map);
  // Morton range used by the count: a lower corner that wrapped negative (a truncated u64 cell id)
  // still admits every cell >= 0, so its code is taken at 0.
  bool wrapped = false;
  for (int i = 0; i < l; ++i) {
    wrapped = wrapped || lo_c[i] < 0;
    lo_c[i] = lo_c[i] < 0 ? 0 : lo_c[i];
  }
  zlo_cnt[qi] = wrapped ? morton_code(lo_c, l, D.max_bits, D.valid_map) : za;
}

// ------------------------------------------------------------------------------------------------
// A3/A4 (sparse).  The integer the reference obtains per (query, sub-cluster) by 2^l-corner
// inclusion-exclusion over compressed prefix sums (Super-Grid-IVF_PQ.cu:648-790) is the number of the
// sub-cluster's points whose cells lie in the query's box.  A selective box holds few points, so they
// are enumerated through the bucket grid: the box covers a block of buckets whose rows along the last
// valid dim are contiguous entry ranges; every entry in them is box-tested on its cells and counted
// into a per-wave LDS histogram over sub-clusters.  The non-zero sub-clusters are then emitted in
// ascending table order with their counts, and the non-zero clusters (filter.cu:242-260 over
// get_cluster_sum_kernel's sums) fall out as the runs of cluster_of_sub.  A query whose box spans more
// than kMaxCombos bucket rows or max_tests entries is handed to the dense count instead.
// One 256-thread workgroup per query (persistent over queries): the four waves share the query's walk and
// one LDS histogram.  LDS = hist[CS] + bit set[CS/32] + word prefixes[CS/32].
struct Buckets {
  uint32_t M[kMaxDims], nb[kMaxDims];
  uint64_t stride[kMaxDims];
  uint32_t cells16;   // entry cells are u16 (else u32)
  uint32_t packed;    // esub holds sub << pk_shift | (position - seg_off[sub]); epos is unused
  uint32_t pk_shift;
};

// Cell of valid dim j of bucket entry e.
__device__ __forceinline__ uint32_t entry_cell(const Buckets& B, const uint32_t* __restrict__ ecell, uint64_t N,
                                               int j, uint32_t e) {
  return B.cells16 ? (uint32_t)reinterpret_cast<const uint16_t*>(ecell)[(uint64_t)j * N + e]
                   : ecell[(uint64_t)j * N + e];
}

constexpr uint64_t kMaxCombos = 4096;
constexpr uint32_t kSparseMaxCS = 8192;  // LDS histogram: (CS + 2 CS/32) words per workgroup

template <int NV>
__global__ __launch_bounds__(256) void sparse_count_kernel(
    Buckets B, const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ ecell,
    const uint32_t* __restrict__ esub, uint64_t N, const uint32_t* __restrict__ sub_off, uint32_t CS,
    uint32_t C, const uint32_t* __restrict__ box_lo, const uint32_t* __restrict__ box_w, uint64_t qn, uint64_t G,
    uint64_t q0, uint32_t max_tests, const uint32_t* __restrict__ epos, uint32_t ibcap,
    uint32_t* __restrict__ ib_pos, uint32_t* __restrict__ ib_sub, uint32_t* __restrict__ n_ib,
    uint32_t* __restrict__ sub_list, uint32_t* __restrict__ sub_cnt,
    uint32_t* __restrict__ ns_out, uint32_t* __restrict__ list_c, uint32_t* __restrict__ cl_beg,
    uint32_t* __restrict__ nnz_c, uint32_t* __restrict__ dense_list, uint32_t* __restrict__ n_dense,
    unsigned long long* __restrict__ gstat, uint32_t* __restrict__ nzb, uint32_t* __restrict__ nzp,
    const uint32_t* __restrict__ qlist, const uint32_t* __restrict__ qlist_n,
    const uint32_t* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  extern __shared__ __align__(16) uint32_t sm[];
  const uint32_t nw = (CS + 31) >> 5;
  uint32_t* hist = sm;            // [CS]
  uint32_t* bits = sm + CS;       // [nw]
  uint32_t* wpre = bits + nw;     // [nw] rank of each word's first set bit among the non-zero sub-clusters
  __shared__ uint32_t pre_sh[kWG], e0_sh[kWG], tmp[kWG / 64 + 1];
  const int tid = threadIdx.x;
  for (uint32_t i = tid; i < CS; i += kWG) hist[i] = 0;
  for (uint32_t i = tid; i < nw; i += kWG) bits[i] = 0;
  __syncthreads();
  // all queries, or only those listed in qlist[0 .. *qlist_n)
  const uint64_t nslots = qlist ? (uint64_t)*qlist_n : qn;
  for (uint64_t slot = blockIdx.x; slot < nslots; slot += gridDim.x) {
    const uint64_t qi = qlist ? (uint64_t)qlist[slot] : slot;
    uint32_t lo[NV], w[NV], blo[NV], bhi[NV];
    bool empty = false;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      lo[j] = box_lo[(uint64_t)j * qn + qi];
      w[j] = box_w[(uint64_t)j * qn + qi];
      // the box test (cell - lo) mod 2^32 <= w admits exactly the cells in [max(lo, 0), lo + w]
      const int lw = (int)lo[j], h = (int)(lo[j] + w[j]);
      const int a = lw < 0 ? 0 : lw;
      if (h < a) empty = true;
      blo[j] = h < a ? 0u : (uint32_t)((uint64_t)a * B.nb[j] / B.M[j]);
      bhi[j] = h < a ? 0u : (uint32_t)((uint64_t)h * B.nb[j] / B.M[j]);
    }
    uint64_t combos = 1;
#pragma unroll
    for (int j = 0; j + 1 < NV; ++j) combos *= (uint64_t)(bhi[j] - blo[j] + 1);
    bool to_dense = !empty && combos > kMaxCombos;
    uint32_t ibn = 0;  // in-box points listed so far (order: pass, thread, entry of the thread)
    if (!empty && !to_dense) {
      uint64_t tests = 0;
      for (uint64_t c0 = 0; c0 < combos; c0 += kWG) {
        const uint64_t c = c0 + (uint64_t)tid;
        uint32_t e0 = 0, len = 0;
        if (c < combos) {
          uint64_t rem = c, base = 0;
#pragma unroll
          for (int j = NV - 2; j >= 0; --j) {
            const uint64_t wj = bhi[j] - blo[j] + 1;
            base += (uint64_t)(blo[j] + (uint32_t)(rem % wj)) * B.stride[j];
            rem /= wj;
          }
          e0 = bstart[base + blo[NV - 1]];
          len = bstart[base + bhi[NV - 1] + 1] - e0;
        }
        uint32_t total;
        const uint32_t excl = dev::wg_exclusive_scan(len, tmp, total);
        pre_sh[tid] = excl;
        e0_sh[tid] = e0;
        __syncthreads();
        tests += total;
        if (tests > max_tests) {
          to_dense = true;
          break;
        }
        // four entries per thread per pass, so their loads are in flight together
        for (uint32_t t0 = 0; t0 < total; t0 += 4 * kWG) {
          uint32_t ev[4], cv[4][NV], sv[4], pv[4];
          // last k with pre_sh[k] <= t (that thread's range holds t): the four searches step together
          static_assert(kWG == 256, "eight search steps");
          uint32_t a[4] = {0u, 0u, 0u, 0u};
#pragma unroll
          for (uint32_t step = kWG / 2; step > 0; step >>= 1)
#pragma unroll
            for (int u = 0; u < 4; ++u)
              if (pre_sh[a[u] + step] <= t0 + (uint32_t)tid + (uint32_t)(kWG * u)) a[u] += step;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint32_t t = t0 + (uint32_t)tid + (uint32_t)(kWG * u);
            ev[u] = t < total ? e0_sh[a[u]] + (t - pre_sh[a[u]]) : kNone;
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const bool ok = ev[u] != kNone;
#pragma unroll
            for (int j = 0; j < NV; ++j) cv[u][j] = ok ? entry_cell(B, ecell, N, j, ev[u]) : lo[j] - 1u;
            const uint32_t se = ok ? esub[ev[u]] : 0u;
            sv[u] = B.packed ? se >> B.pk_shift : se;
            pv[u] = B.packed ? se : (ok ? epos[ev[u]] : 0u);
          }
          bool in[4];
          uint32_t nin = 0;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            in[u] = ev[u] != kNone;
#pragma unroll
            for (int j = 0; j < NV; ++j) in[u] = in[u] && (cv[u][j] - lo[j] <= w[j]);
            nin += in[u] ? 1u : 0u;
          }
          uint32_t pass_tot;
          uint32_t at = ibn + dev::wg_exclusive_scan(nin, tmp, pass_tot);
          ibn += pass_tot;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            if (in[u]) {
              const uint32_t s2 = sv[u];
              if (atomicAdd(&hist[s2], 1u) == 0u) atomicOr(&bits[s2 >> 5], 1u << (s2 & 31));
              if (at < ibcap) {
                ib_pos[qi * ibcap + at] = pv[u];
                if (!B.packed) ib_sub[qi * ibcap + at] = s2;
              }
              ++at;
            }
          }
        }
        __syncthreads();  // pre_sh / e0_sh are rewritten by the next batch of bucket rows
      }
    }
    __syncthreads();
    if (to_dense) {
      for (uint32_t wi = tid; wi < nw; wi += kWG) {
        uint32_t word = bits[wi];
        while (word) {
          const uint32_t b = __ffs(word) - 1;
          word &= word - 1;
          hist[wi * 32 + b] = 0;
        }
        bits[wi] = 0;
      }
      if (tid == 0) {
        dense_list[atomicAdd(n_dense, 1u)] = (uint32_t)qi;
        n_ib[qi] = kNone;
      }
      __syncthreads();
      continue;
    }
    // emit the non-zero sub-clusters ascending with their counts
    uint32_t* sl = sub_list + qi * CS;
    uint32_t* sc = sub_cnt + qi * CS;
    uint32_t base = 0;
    for (uint32_t w0 = 0; w0 < nw; w0 += kWG) {
      const uint32_t wi = w0 + (uint32_t)tid;
      const uint32_t word = wi < nw ? bits[wi] : 0u;
      uint32_t tot;
      uint32_t pos = base + dev::wg_exclusive_scan(__popc(word), tmp, tot);
      if (wi < nw) wpre[wi] = pos;
      uint32_t rest = word;
      while (rest) {
        const uint32_t b = __ffs(rest) - 1;
        rest &= rest - 1;
        const uint32_t s2 = wi * 32 + b;
        sl[pos] = s2;
        sc[pos] = hist[s2];
        hist[s2] = 0;
        ++pos;
      }
      base += tot;
    }
    __syncthreads();
    // the non-zero clusters, ascending, each with the rank of its first non-zero sub-cluster (sub-clusters
    // are numbered cluster-major, so cluster c owns [sub_off[c], sub_off[c+1]))
    uint32_t* lc = list_c + qi * C;
    uint32_t* cb = cl_beg + qi * (C + 1);
    uint32_t ncl = 0;
    for (uint32_t c0 = 0; c0 < C; c0 += kWG) {
      const uint32_t c = c0 + (uint32_t)tid;
      uint32_t first = kNone;
      if (c < C) {
        const uint32_t a = sub_off[c], e = sub_off[c + 1];
        for (uint32_t x = a; x < e && first == kNone;) {
          const uint32_t wi = x >> 5, sh = x & 31;
          const uint32_t span = e - x < 32 - sh ? e - x : 32 - sh;
          const uint32_t m = (bits[wi] >> sh) & (span == 32 ? 0xFFFFFFFFu : ((1u << span) - 1u));
          if (m) first = x + (uint32_t)(__ffs(m) - 1);
          x += span;
        }
      }
      const uint32_t flag = first != kNone ? 1u : 0u;
      uint32_t tot;
      const uint32_t at = ncl + dev::wg_exclusive_scan(flag, tmp, tot);
      if (flag) {
        const uint32_t wi = first >> 5, sh = first & 31;
        lc[at] = c;
        cb[at] = wpre[wi] + (uint32_t)__popc(bits[wi] & ((1u << sh) - 1u));
      }
      ncl += tot;
    }
    __syncthreads();
    for (uint32_t wi = tid; wi < nw; wi += kWG) {
      if (nzb) {
        nzb[qi * nw + wi] = bits[wi];
        nzp[qi * nw + wi] = wpre[wi];
      }
      bits[wi] = 0;
    }
    if (tid == 0) {
      ns_out[qi] = base;
      nnz_c[qi] = ncl;
      cb[ncl] = base;
      n_ib[qi] = ibn <= ibcap ? ibn : kNone;
      atomicMax(&gstat[((q0 + qi) / G) * 8 + 0], (unsigned long long)ncl);
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------------------
// A3/A4 (sparse), one wave per query: the same walk, in-box list, counts and lists as sparse_count_kernel,
// without its LDS histogram and barriers.  The wave keeps the sub-cluster ids of its in-box points in a wave
// buffer of kWaveIb entries, sorts them in registers, and takes the non-zero sub-clusters (ascending, with
// exact counts) from their runs; an LDS bit set of the distinct ids with word ranks gives the non-zero
// clusters (cluster c owns sub-clusters [sub_off[c], sub_off[c+1])) and the bit words + ranks.  A query with more in-box points
// than kWaveIb goes on `redo` for sparse_count_kernel; a wide box goes to the dense count as there.
constexpr uint32_t kWaveIb = 512;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  return x;
}

__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

template <int R>
__device__ void wave_sort_buf(uint32_t* buf, uint32_t n, int lane) {
  uint32_t key[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)lane * R + r;
    key[r] = e < n ? buf[e] : 0xFFFFFFFFu;
  }
  dev::wave_bitonic32<R>(key, lane);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)lane * R + r;
    if (e < n) buf[e] = key[r];
  }
}

template <int NV>
__global__ __launch_bounds__(256) void sparse_count_wave_kernel(
    Buckets B, const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ ecell,
    const uint32_t* __restrict__ esub, uint64_t N, const uint32_t* __restrict__ sub_off, uint32_t CS,
    uint32_t C, const uint32_t* __restrict__ box_lo, const uint32_t* __restrict__ box_w, uint64_t qn, uint64_t G,
    uint64_t q0, uint32_t max_tests, const uint32_t* __restrict__ epos, uint32_t ibcap,
    uint32_t* __restrict__ ib_pos, uint32_t* __restrict__ ib_sub, uint32_t* __restrict__ n_ib,
    uint32_t* __restrict__ sub_list, uint32_t* __restrict__ sub_cnt,
    uint32_t* __restrict__ ns_out, uint32_t* __restrict__ list_c, uint32_t* __restrict__ cl_beg,
    uint32_t* __restrict__ nnz_c, uint32_t* __restrict__ dense_list, uint32_t* __restrict__ n_dense,
    unsigned long long* __restrict__ gstat, uint32_t* __restrict__ nzb, uint32_t* __restrict__ nzp,
    uint32_t* __restrict__ redo, uint32_t skip, uint32_t* __restrict__ abort_w) {
  __shared__ uint32_t sbuf_all[4][kWaveIb];  // in-box sub-cluster ids; sorted; then first positions of runs
  __shared__ uint32_t ubuf_all[4][kWaveIb];  // unique sub-cluster ids in rank order
  __shared__ uint32_t words_all[4][kSparseMaxCS / 32], wrank_all[4][kSparseMaxCS / 32];  // bit set, word ranks
  __shared__ uint32_t pre_all[4][64], e0_all[4][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t* sbuf = sbuf_all[wv];
  uint32_t* ubuf = ubuf_all[wv];
  uint32_t* words = words_all[wv];
  uint32_t* wrank = wrank_all[wv];
  uint32_t* pre_sh = pre_all[wv];
  uint32_t* e0_sh = e0_all[wv];
  const uint32_t nw = (CS + 31) >> 5;
  const uint64_t lt = (1ull << lane) - 1ull;
  for (uint64_t qi = (uint64_t)blockIdx.x * 4 + wv; qi < qn; qi += (uint64_t)gridDim.x * 4) {
    uint32_t lo[NV], w[NV], blo[NV], bhi[NV];
    bool empty = false;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      lo[j] = box_lo[(uint64_t)j * qn + qi];
      w[j] = box_w[(uint64_t)j * qn + qi];
      // the box test (cell - lo) mod 2^32 <= w admits exactly the cells in [max(lo, 0), lo + w]
      const int lw = (int)lo[j], h = (int)(lo[j] + w[j]);
      const int a = lw < 0 ? 0 : lw;
      if (h < a) empty = true;
      blo[j] = h < a ? 0u : (uint32_t)((uint64_t)a * B.nb[j] / B.M[j]);
      bhi[j] = h < a ? 0u : (uint32_t)((uint64_t)h * B.nb[j] / B.M[j]);
    }
    uint64_t combos = 1;
#pragma unroll
    for (int j = 0; j + 1 < NV; ++j) combos *= (uint64_t)(bhi[j] - blo[j] + 1);
    bool to_dense = !empty && combos > kMaxCombos;
    uint32_t ibn = 0;  // in-box points listed so far (order: bucket-row batch, entry slot, lane)
    if (!empty && !to_dense) {
      uint64_t tests = 0;
      for (uint64_t c0 = 0; c0 < combos; c0 += 64) {
        const uint64_t c = c0 + (uint64_t)lane;
        uint32_t e0 = 0, len = 0;
        if (c < combos) {
          uint64_t rem = c, base = 0;
#pragma unroll
          for (int j = NV - 2; j >= 0; --j) {
            const uint64_t wj = bhi[j] - blo[j] + 1;
            base += (uint64_t)(blo[j] + (uint32_t)(rem % wj)) * B.stride[j];
            rem /= wj;
          }
          e0 = bstart[base + blo[NV - 1]];
          len = bstart[base + bhi[NV - 1] + 1] - e0;
        }
        const uint32_t incl = wave_incl_scan(len, lane);
        const uint32_t total = (uint32_t)__shfl((int)incl, 63, 64);
        pre_sh[lane] = incl - len;
        e0_sh[lane] = e0;
        wave_lds_fence();
        tests += total;
        if (tests > max_tests) {
          to_dense = true;
          break;
        }
        // four entries per lane per pass, so their loads are in flight together.  Once the in-box points
        // overflow the wave buffer the query goes to the workgroup walk, which walks the box again: the
        // entries are not tested further, but the remaining bucket rows are still summed, so that a box
        // over max_tests entries goes to the dense count as before
        for (uint32_t t0 = 0; t0 < total && ibn <= kWaveIb; t0 += 256) {
          uint32_t ev[4], cv[4][NV], sv[4], pv[4];
          // last k with pre_sh[k] <= t (that lane's range holds t): the four searches step together
          uint32_t a[4] = {0u, 0u, 0u, 0u};
#pragma unroll
          for (uint32_t step = 32; step > 0; step >>= 1)
#pragma unroll
            for (int u = 0; u < 4; ++u)
              if (pre_sh[a[u] + step] <= t0 + (uint32_t)lane + 64u * u) a[u] += step;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint32_t t = t0 + (uint32_t)lane + 64u * u;
            ev[u] = t < total ? e0_sh[a[u]] + (t - pre_sh[a[u]]) : kNone;
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const bool ok = ev[u] != kNone;
#pragma unroll
            for (int j = 0; j < NV; ++j) cv[u][j] = ok ? entry_cell(B, ecell, N, j, ev[u]) : lo[j] - 1u;
            const uint32_t se = ok ? esub[ev[u]] : 0u;
            sv[u] = B.packed ? se >> B.pk_shift : se;
            pv[u] = B.packed ? se : (ok ? epos[ev[u]] : 0u);
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            bool in = ev[u] != kNone;
#pragma unroll
            for (int j = 0; j < NV; ++j) in = in && (cv[u][j] - lo[j] <= w[j]);
            const uint64_t mask = __ballot(in);
            const uint32_t at = ibn + (uint32_t)__popcll(mask & lt);
            ibn += (uint32_t)__popcll(mask);
            if (in) {
              if (at < ibcap) {
                ib_pos[qi * ibcap + at] = pv[u];
                if (!B.packed) ib_sub[qi * ibcap + at] = sv[u];
              }
              if (at < kWaveIb) sbuf[at] = sv[u];
            }
          }
        }
        wave_lds_fence();  // pre_sh / e0_sh are rewritten by the next batch of bucket rows
      }
    }
    // hand-overs; a replayed graph that left the receiving kernels out is abandoned (skip bits 1 / 2)
    if (to_dense) {
      if (lane == 0) {
        dense_list[atomicAdd(&n_dense[0], 1u)] = (uint32_t)qi;
        n_ib[qi] = kNone;
        if (skip & 1u) *abort_w = 1u;
      }
      continue;
    }
    if (ibn > kWaveIb) {
      if (lane == 0) {
        redo[atomicAdd(&n_dense[1], 1u)] = (uint32_t)qi;
        if (skip & 2u) *abort_w = 1u;
      }
      continue;
    }
    wave_lds_fence();
    const uint32_t n = ibn;
    if (n > 1) {
      if (n <= 64) wave_sort_buf<1>(sbuf, n, lane);
      else if (n <= 128) wave_sort_buf<2>(sbuf, n, lane);
      else if (n <= 256) wave_sort_buf<4>(sbuf, n, lane);
      else wave_sort_buf<8>(sbuf, n, lane);
      wave_lds_fence();
    }
    // runs of the sorted ids: ubuf[u] = the u-th distinct id, sbuf[u] = its first position (written in
    // place: position u <= e, and a batch reads its elements before writing)
    uint32_t nu = 0, carry = 0xFFFFFFFFu;
    for (uint32_t b0 = 0; b0 < n; b0 += 64) {
      const uint32_t e = b0 + (uint32_t)lane;
      const uint32_t v = e < n ? sbuf[e] : 0xFFFFFFFFu;
      uint32_t prev = (uint32_t)__shfl_up((int)v, 1, 64);
      if (lane == 0) prev = carry;
      const bool start = e < n && (e == 0 || v != prev);
      carry = (uint32_t)__shfl((int)v, 63, 64);
      const uint64_t mask = __ballot(start);
      const uint32_t u = nu + (uint32_t)__popcll(mask & lt);
      wave_lds_fence();
      if (start) {
        ubuf[u] = v;
        sbuf[u] = e;
      }
      nu += (uint32_t)__popcll(mask);
      wave_lds_fence();
    }
    uint32_t* sl = sub_list + qi * CS;
    uint32_t* sc = sub_cnt + qi * CS;
    for (uint32_t u = lane; u < nu; u += 64) {
      sl[u] = ubuf[u];
      sc[u] = (u + 1 < nu ? sbuf[u + 1] : n) - sbuf[u];
    }
    // bit set of the distinct ids and the rank of each word's first one
    for (uint32_t wi = lane; wi < nw; wi += 64) words[wi] = 0;
    wave_lds_fence();
    for (uint32_t u = lane; u < nu; u += 64) atomicOr(&words[ubuf[u] >> 5], 1u << (ubuf[u] & 31));
    wave_lds_fence();
    uint32_t wbase = 0;
    for (uint32_t w0 = 0; w0 < nw; w0 += 64) {
      const uint32_t wi = w0 + (uint32_t)lane;
      const uint32_t word = wi < nw ? words[wi] : 0u;
      const uint32_t pc = (uint32_t)__popc(word);
      const uint32_t incl = wave_incl_scan(pc, lane);
      if (wi < nw) {
        wrank[wi] = wbase + incl - pc;
        if (nzb) {
          nzb[qi * nw + wi] = word;
          nzp[qi * nw + wi] = wbase + incl - pc;
        }
      }
      wbase += (uint32_t)__shfl((int)incl, 63, 64);
    }
    wave_lds_fence();
    // the non-zero clusters, ascending, each with the rank of its first non-zero sub-cluster
    uint32_t* lc = list_c + qi * C;
    uint32_t* cb = cl_beg + qi * (C + 1);
    uint32_t ncl = 0;
    for (uint32_t c0 = 0; c0 < C; c0 += 64) {
      const uint32_t c = c0 + (uint32_t)lane;
      uint32_t first = kNone;
      if (c < C && nu > 0) {
        const uint32_t a = sub_off[c], e = sub_off[c + 1];
        for (uint32_t x = a; x < e && first == kNone;) {
          const uint32_t wi = x >> 5, sh = x & 31;
          const uint32_t span = e - x < 32 - sh ? e - x : 32 - sh;
          const uint32_t m = (words[wi] >> sh) & (span == 32 ? 0xFFFFFFFFu : ((1u << span) - 1u));
          if (m) first = x + (uint32_t)(__ffs(m) - 1);
          x += span;
        }
      }
      const bool nz = first != kNone;
      const uint64_t mask = __ballot(nz);
      if (nz) {
        const uint32_t at = ncl + (uint32_t)__popcll(mask & lt);
        const uint32_t wi = first >> 5, sh = first & 31;
        lc[at] = c;
        cb[at] = wrank[wi] + (uint32_t)__popc(words[wi] & ((1u << sh) - 1u));
      }
      ncl += (uint32_t)__popcll(mask);
    }
    if (lane == 0) {
      ns_out[qi] = nu;
      nnz_c[qi] = ncl;
      cb[ncl] = nu;
      n_ib[qi] = ibn;
      atomicMax(&gstat[((q0 + qi) / G) * 8 + 0], (unsigned long long)ncl);
    }
    wave_lds_fence();  // the buffers are reused by the wave's next query
  }
}

// ------------------------------------------------------------------------------------------------
// A3/A4 (dense), for the queries the sparse pass handed over (wide boxes): which sub-clusters hold at
// least one in-box point.  A point inside the box has a Morton code in [z(lo), z(hi)] (the bit interleave
// is monotone in every coordinate) and a sub-cluster's points are sorted by Morton code, so only that code
// range of each segment is box-tested, and a lane stops at its first hit.  The exact in-box counts are
// needed only for the chosen sub-clusters (A10's beta) and are taken there (ranges_kernel).  Block = 256
// lanes = 256 handed-over queries; blockIdx.y = chunk of kCountSegChunk segments.  Each segment is
// staged in LDS in pieces of kCountPiece points (codes + cells); every lane binary-searches its own
// code range in LDS.  counts[slot][s] = 0 / 1, slot = position in dense_list.
constexpr int kCountPiece = 1024;

// Coarse occupancy sums of the index (per sub-cluster inclusive prefix sums over coarse cells, shifted by one).
struct CoarseSums {
  const uint32_t* t = nullptr;  // [CS][per], or null: no table
  uint64_t per = 0;
  uint32_t sg[kMaxDims] = {0}, Mv[kMaxDims] = {0};
  uint64_t stride[kMaxDims] = {0};
};

// Points of a sub-cluster whose coarse cells lie in [a_j, b_j] in every dim: inclusion-exclusion over the
// 2^NV corners of its prefix-sum table t.
template <int NV>
__device__ __forceinline__ uint32_t coarse_rect(const uint32_t* __restrict__ t, const uint64_t* stride,
                                                const uint32_t* a, const uint32_t* b) {
  int64_t sum = 0;
#pragma unroll
  for (uint32_t m = 0; m < (1u << NV); ++m) {
    uint64_t x = 0;
    int low = 0;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      if ((m >> j) & 1u) {
        x += ((uint64_t)b[j] + 1) * stride[j];
      } else {
        x += (uint64_t)a[j] * stride[j];
        ++low;
      }
    }
    sum += (low & 1) ? -(int64_t)t[x] : (int64_t)t[x];
  }
  return (uint32_t)sum;
}

template <int NV>
__global__ __launch_bounds__(256) void dense_count_kernel(const uint32_t* __restrict__ cells, uint64_t N,
                                                          const uint64_t* __restrict__ zcode,
                                                          const uint32_t* __restrict__ seg_off, uint64_t CS,
                                                          const uint32_t* __restrict__ box_lo,
                                                          const uint32_t* __restrict__ box_w,
                                                          const uint64_t* __restrict__ zlo,
                                                          const uint64_t* __restrict__ zhi, uint64_t qn,
                                                          const uint32_t* __restrict__ dense_list,
                                                          const uint32_t* __restrict__ n_dense,
                                                          uint32_t* __restrict__ counts,
    const uint32_t* __restrict__ abort_flag, CoarseSums sums) {
  if (abort_flag && *abort_flag) return;
  __shared__ uint64_t zs[kCountPiece];
  __shared__ uint32_t cs[NV][kCountPiece];
  const uint32_t nd = *n_dense;
  const uint64_t slot = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if ((uint64_t)blockIdx.x * 256 >= nd) return;  // uniform per block
  const bool live = slot < nd;
  const uint64_t q = live ? dense_list[slot] : 0;
  uint32_t lo[NV], w[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    lo[j] = live ? box_lo[(uint64_t)j * qn + q] : 1u;
    w[j] = live ? box_w[(uint64_t)j * qn + q] : 0u;
  }
  const uint64_t za = live ? zlo[q] : ~0ull, zb = live ? zhi[q] : 0ull;
  // the box in coarse cells: every coarse cell it touches (o) and the ones it covers whole (i)
  uint32_t oa[NV], ob[NV], ia[NV], ib[NV];
  bool interior = true, none = !live;
  if (sums.t) {
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const int lw = (int)lo[j], h = (int)(lo[j] + w[j]);
      const int a = lw < 0 ? 0 : lw;  // the box test admits the cells [max(lo, 0), lo + w]
      if (h < a) none = true;
      const uint64_t G = sums.sg[j], M = sums.Mv[j];
      oa[j] = (uint32_t)((uint64_t)(uint32_t)a * G / M);
      ob[j] = h < a ? oa[j] : (uint32_t)((uint64_t)(uint32_t)h * G / M);
      // coarse cell g holds the cells [ceil(g M / G), ceil((g + 1) M / G) - 1]
      const uint64_t first = ((uint64_t)oa[j] * M + G - 1) / G, last = (((uint64_t)ob[j] + 1) * M + G - 1) / G - 1;
      ia[j] = first >= (uint64_t)a ? oa[j] : oa[j] + 1;
      const int64_t ibj = last <= (uint64_t)(int64_t)h ? (int64_t)ob[j] : (int64_t)ob[j] - 1;
      if (ibj < (int64_t)ia[j]) interior = false;
      ib[j] = ibj < 0 ? 0u : (uint32_t)ibj;
    }
  }
  const uint64_t s_begin = (uint64_t)blockIdx.y * kCountSegChunk;
  const uint64_t s_end = s_begin + kCountSegChunk < CS ? s_begin + kCountSegChunk : CS;
  for (uint64_t s = s_begin; s < s_end; ++s) {
    const uint32_t p0 = seg_off[s], p1 = seg_off[s + 1];
    bool hit = !live;    // idle lanes never hold the block back
    bool done = none;    // settled without a hit
    if (sums.t && live && !none) {
      // a point in a covered coarse cell is in the box; no point in a touched one: none is
      const uint32_t* t = sums.t + s * sums.per;
      if (interior && coarse_rect<NV>(t, sums.stride, ia, ib) > 0) hit = true;
      else if (coarse_rect<NV>(t, sums.stride, oa, ob) == 0) done = true;
    }
    for (uint32_t base = p0; base < p1; base += kCountPiece) {
      const uint32_t m = p1 - base < (uint32_t)kCountPiece ? p1 - base : (uint32_t)kCountPiece;
      if (__syncthreads_count(hit || done ? 0 : 1) == 0) break;  // every lane has its answer
      for (uint32_t i = threadIdx.x; i < m; i += 256) {
        zs[i] = zcode[base + i];
#pragma unroll
        for (int j = 0; j < NV; ++j) cs[j][i] = cells[(uint64_t)j * N + base + i];
      }
      __syncthreads();
      if (hit || done || za > zs[m - 1] || zb < zs[0]) continue;
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
        if (in) {
          hit = true;
          break;
        }
      }
    }
    if (live) counts[slot * CS + s] = hit ? 1u : 0u;
  }
}

// Non-zero sub-clusters (ascending, with counts) and non-zero clusters of a densely counted query.
// Phase 1 overwrites each count with the number of non-zero sub-clusters before it, which phase 2
// reads at the cluster boundaries.  One workgroup per handed-over query.
__global__ __launch_bounds__(256) void dense_emit_kernel(uint32_t* __restrict__ counts, uint32_t CS,
                                                         const uint32_t* __restrict__ sub_off, uint32_t C,
                                                         uint64_t G, uint64_t q0,
                                                         const uint32_t* __restrict__ dense_list,
                                                         const uint32_t* __restrict__ n_dense,
                                                         uint32_t* __restrict__ sub_list, uint32_t* __restrict__ sub_cnt,
                                                         uint32_t* __restrict__ ns_out, uint32_t* __restrict__ list_c,
                                                         uint32_t* __restrict__ cl_beg, uint32_t* __restrict__ nnz_c,
                                                         unsigned long long* __restrict__ gstat,
                                                         uint32_t* __restrict__ nzb, uint32_t* __restrict__ nzp,
    const uint32_t* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  __shared__ uint32_t tmp[8];
  const uint64_t slot = blockIdx.x;
  if (slot >= *n_dense) return;
  const uint64_t q = dense_list[slot];
  uint32_t* row = counts + slot * CS;
  uint32_t base = 0;
  for (uint32_t b = 0; b < CS; b += kWG) {
    const uint32_t s = b + threadIdx.x;
    const uint32_t c = s < CS ? row[s] : 0u;
    uint32_t total;
    const uint32_t pos = dev::wg_exclusive_scan(c != 0 ? 1u : 0u, tmp, total);
    if (c != 0) {
      sub_list[q * CS + base + pos] = s;
      sub_cnt[q * CS + base + pos] = c;
    }
    if (nzb) {
      // bit words of the non-zero sub-clusters and the rank of each word's first one
      const uint64_t bal = __ballot(c != 0);
      const uint32_t lane = threadIdx.x & 63u, nw = (CS + 31) >> 5, wi = s >> 5;
      if ((lane == 0 || lane == 32) && wi < nw) {
        nzb[q * nw + wi] = lane == 0 ? (uint32_t)bal : (uint32_t)(bal >> 32);
        nzp[q * nw + wi] = base + pos;
      }
    }
    if (s < CS) row[s] = base + pos;
    base += total;
  }
  __syncthreads();
  const uint32_t ns = base;
  uint32_t nc = 0;
  for (uint32_t b = 0; b < C; b += kWG) {
    const uint32_t c = b + threadIdx.x;
    uint32_t beg = 0, end = 0;
    if (c < C) {
      const uint32_t s0 = sub_off[c], s1 = sub_off[c + 1];
      beg = s0 < CS ? row[s0] : ns;
      end = s1 < CS ? row[s1] : ns;
    }
    const uint32_t flag = end > beg ? 1u : 0u;
    uint32_t total;
    const uint32_t pos = dev::wg_exclusive_scan(flag, tmp, total);
    if (flag) {
      list_c[q * C + nc + pos] = c;
      cl_beg[q * (C + 1) + nc + pos] = beg;
    }
    nc += total;
  }
  if (threadIdx.x == 0) {
    ns_out[q] = ns;
    nnz_c[q] = nc;
    cl_beg[q * (C + 1) + nc] = ns;
    atomicMax(&gstat[((q0 + q) / G) * 8 + 0], (unsigned long long)nc);
  }
}

// ------------------------------------------------------------------------------------------------
// Exact FP32 L2 of listed rows, the reference's order for d <= 512 (refine.cu:36-75): a sequential f32
// sum over dims of rounded squares, one lane per row.  Work = (query, 64 consecutive list entries);
// woff[q] = chunks before query q (scan_local_kernel), woff[qn] = total.  One wave per workgroup; each
// workgroup takes a contiguous run of chunks, so the query vector is staged once per query.  Rows pass
// through LDS 64 dims at a time with 16 lanes per row (each wave load reads four whole 256-byte row
// pieces); the next stage's loads are in flight while the current stage is summed from LDS.
// KIND 0: rows = list[q * lstride + j], key tie = j.  KIND 1: rows = list[loff[q] + j] (z-pool
// positions), key tie = zid[row].
constexpr int kDistStage = 64;
constexpr int kStageLd = 68;  // LDS row stride in floats (16-byte aligned rows)

__device__ __forceinline__ uint64_t owner_of(const uint64_t* __restrict__ woff, uint64_t qn, uint64_t k) {
  uint64_t a = 0, b = qn;  // last q with woff[q] <= k; that query owns chunk k
  while (b - a > 1) {
    const uint64_t m = (a + b) >> 1;
    if (woff[m] <= k) a = m; else b = m;
  }
  return a;
}

__device__ __forceinline__ float sq_chain4(float acc, float4 qq, float4 x) {
  float diff = qq.x - x.x;
  float sq = diff * diff;
  acc = acc + sq;
  diff = qq.y - x.y;
  sq = diff * diff;
  acc = acc + sq;
  diff = qq.z - x.z;
  sq = diff * diff;
  acc = acc + sq;
  diff = qq.w - x.w;
  sq = diff * diff;
  acc = acc + sq;
  return acc;
}

// Rows of at least this many dims are streamed past the caches by the exact re-rank (measured: 1,536-byte rows
// of the 1M x 384 workload gain, 512-byte rows of the 5M x 128 one, which neighbouring queries share, lose).
constexpr uint64_t kStreamRowsMinDim = 256;
// Largest vector dimension: the WIDE exact-distance kernel holds the query (padded to 512) in 32 KiB of LDS.
constexpr uint64_t kMaxDim = 8192;

// WIDE (d > 512): refine.cu runs min(d, 512) = 512 threads; thread t's sum over dims t, t + 512, ... is an fma
// chain (nvcc contracts `sum += diff * diff`), and thread 0 adds the 512 partials in order.  The lane keeps the
// 64 partials of one block of t in registers while the stages for t0 + 512 c (c = 0, 1, ...) pass through LDS,
// then adds them to its total; the query vector sits in dynamic LDS, zero padded to a multiple of 512.
template <int KIND, bool STREAM = false, bool WIDE = false>
__global__ __launch_bounds__(64) void pair_dist_kernel(const float* __restrict__ queries, uint64_t q0, uint64_t d,
                                                       const float* __restrict__ X, const uint64_t* __restrict__ woff,
                                                       uint64_t qn, const uint32_t* __restrict__ list,
                                                       uint64_t lstride, const uint64_t* __restrict__ loff,
                                                       const uint32_t* __restrict__ nrows,
                                                       const uint32_t* __restrict__ zid, uint64_t* __restrict__ keys,
                                                       const uint32_t* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  __shared__ __align__(16) float qv_s[WIDE ? 4 : 512];
  extern __shared__ __align__(16) float qv_dyn[];
  float* const qv = WIDE ? qv_dyn : qv_s;
  __shared__ __align__(16) float stage[64 * kStageLd];
  __shared__ uint32_t rows[64];
  const int lane = threadIdx.x;
  const uint64_t total = woff[qn];
  const uint64_t per = (total + gridDim.x - 1) / gridDim.x;
  uint64_t k = (uint64_t)blockIdx.x * per;
  const uint64_t kend = k + per < total ? k + per : total;
  if (k >= kend) return;
  uint64_t q = owner_of(woff, qn, k);
  uint64_t cur = ~0ull;
  const bool vec4 = (d & 3) == 0;
  for (; k < kend; ++k) {
    while (woff[q + 1] <= k) ++q;
    const uint64_t chunk = k - woff[q];
    const uint64_t lb = KIND == 1 ? loff[q] : q * lstride;
    const uint32_t n = nrows[q];
    if (q != cur) {
      const uint64_t qlen = WIDE ? (d + 511) / 512 * 512 : d;
      for (uint64_t t = lane; t < qlen; t += 64) qv[t] = t < d ? queries[(q0 + q) * d + t] : 0.0f;
      cur = q;
    }
    const uint32_t j = (uint32_t)chunk * 64 + lane;
    const uint32_t myrow = j < n ? list[lb + j] : kNone;
    rows[lane] = myrow;
    __syncthreads();
    float acc = 0.0f;
    if constexpr (WIDE) {
      const uint32_t nper = (uint32_t)((d + 511) / 512), nst = 8 * nper;
      // stage si covers dims [t0, t0 + 64) with t0 = 512 (si % nper) + 64 (si / nper)
      auto stage_t0 = [&](uint32_t si) -> uint64_t { return (uint64_t)(si % nper) * 512 + (uint64_t)(si / nper) * 64; };
      float4 pre[16];
      auto load_stage = [&](uint64_t t0) {
#pragma unroll
        for (int it = 0; it < 16; ++it) {
          const uint32_t e = lane + it * 64;
          const uint32_t r = e >> 4, c4 = (e & 15) * 4;
          const uint32_t row = rows[r];
          float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
          if (row != kNone) {
            const float* src = X + (uint64_t)row * d + t0 + c4;
            if (vec4) {
              if (t0 + c4 < d) v = *reinterpret_cast<const float4*>(src);
            } else {
              if (t0 + c4 < d) v.x = src[0];
              if (t0 + c4 + 1 < d) v.y = src[1];
              if (t0 + c4 + 2 < d) v.z = src[2];
              if (t0 + c4 + 3 < d) v.w = src[3];
            }
          }
          pre[it] = v;
        }
      };
      float part[64];
      load_stage(stage_t0(0));
      for (uint32_t si = 0; si < nst; ++si) {
        const uint64_t t0 = stage_t0(si);
#pragma unroll
        for (int it = 0; it < 16; ++it) {
          const uint32_t e = lane + it * 64;
          *reinterpret_cast<float4*>(stage + (e >> 4) * kStageLd + (e & 15) * 4) = pre[it];
        }
        __syncthreads();
        if (si + 1 < nst) load_stage(stage_t0(si + 1));
        if (myrow != kNone) {
          const float4* s4 = reinterpret_cast<const float4*>(stage + lane * kStageLd);
          const float4* q4 = reinterpret_cast<const float4*>(qv + t0);
          if (si % nper == 0) {
#pragma unroll
            for (int c = 0; c < 64; ++c) part[c] = 0.0f;
          }
#pragma unroll
          for (int c = 0; c < 16; ++c) {
            const float4 a = q4[c], b = s4[c];
            float df = a.x - b.x;
            part[4 * c] = fmaf(df, df, part[4 * c]);
            df = a.y - b.y;
            part[4 * c + 1] = fmaf(df, df, part[4 * c + 1]);
            df = a.z - b.z;
            part[4 * c + 2] = fmaf(df, df, part[4 * c + 2]);
            df = a.w - b.w;
            part[4 * c + 3] = fmaf(df, df, part[4 * c + 3]);
          }
          if (si % nper == nper - 1) {
#pragma unroll
            for (int c = 0; c < 64; ++c) acc = acc + part[c];
          }
        }
        __syncthreads();
      }
    } else if (vec4) {
      float4 pre[16];
      auto load_stage = [&](uint64_t t0) {
#pragma unroll
        for (int it = 0; it < 16; ++it) {
          const uint32_t e = lane + it * 64;
          const uint32_t r = e >> 4, c4 = (e & 15) * 4;
          const uint32_t row = rows[r];
          // STREAM (re-rank of long rows): data rows are rarely read twice, so they are streamed past the
          // caches; (sub-)centers and short rows that neighbouring queries share stay cached
          typedef float nt4 __attribute__((ext_vector_type(4)));
          nt4 v = {0.0f, 0.0f, 0.0f, 0.0f};
          if (row != kNone && t0 + c4 < d) {
            const nt4* src = reinterpret_cast<const nt4*>(X + (uint64_t)row * d + t0 + c4);
            if constexpr (STREAM) v = __builtin_nontemporal_load(src);
            else v = *src;
          }
          pre[it] = make_float4(v.x, v.y, v.z, v.w);
        }
      };
      load_stage(0);
      for (uint64_t t0 = 0; t0 < d; t0 += kDistStage) {
#pragma unroll
        for (int it = 0; it < 16; ++it) {
          const uint32_t e = lane + it * 64;
          *reinterpret_cast<float4*>(stage + (e >> 4) * kStageLd + (e & 15) * 4) = pre[it];
        }
        __syncthreads();
        if (t0 + kDistStage < d) load_stage(t0 + kDistStage);
        if (myrow != kNone) {
          const float4* s4 = reinterpret_cast<const float4*>(stage + lane * kStageLd);
          const float4* q4 = reinterpret_cast<const float4*>(qv + t0);
          const uint32_t n4 = (uint32_t)((d - t0 < (uint64_t)kDistStage ? d - t0 : (uint64_t)kDistStage) / 4);
          if (n4 == 16) {
#pragma unroll
            for (int c = 0; c < 16; ++c) acc = sq_chain4(acc, q4[c], s4[c]);
          } else {
            for (uint32_t c = 0; c < n4; ++c) acc = sq_chain4(acc, q4[c], s4[c]);
          }
        }
        __syncthreads();
      }
    } else {
      for (uint64_t t0 = 0; t0 < d; t0 += kDistStage) {
        const uint32_t wdt = (uint32_t)(d - t0 < (uint64_t)kDistStage ? d - t0 : (uint64_t)kDistStage);
        for (uint32_t e = lane; e < 64u * kDistStage; e += 64) {
          const uint32_t r = e / kDistStage, c = e % kDistStage;
          const uint32_t row = rows[r];
          stage[r * kStageLd + c] = (row != kNone && c < wdt) ? X[(uint64_t)row * d + t0 + c] : 0.0f;
        }
        __syncthreads();
        if (myrow != kNone) {
          const float* st = stage + lane * kStageLd;
          for (uint32_t c = 0; c < wdt; ++c) {
            const float diff = qv[t0 + c] - st[c];
            const float sq = diff * diff;
            acc = acc + sq;
          }
        }
        __syncthreads();
      }
    }
    if (myrow != kNone) keys[lb + j] = dev::make_key(acc, KIND == 1 ? zid[myrow] : j);
  }
}

// pair_dist_kernel launch: the WIDE form (dynamic LDS for the padded query) when d > 512.
template <int KIND, bool STREAM = false>
void launch_pair_dist(unsigned grid, hipStream_t s, const float* queries, uint64_t q0, uint64_t d, const float* X,
                      const uint64_t* woff, uint64_t qn, const uint32_t* list, uint64_t lstride, const uint64_t* loff,
                      const uint32_t* nrows, const uint32_t* zid, uint64_t* keys, const uint32_t* abort_flag) {
  if (d > 512)
    hipLaunchKernelGGL((pair_dist_kernel<KIND, STREAM, true>), dim3(grid), dim3(64), (d + 511) / 512 * 512 * 4, s,
                       queries, q0, d, X, woff, qn, list, lstride, loff, nrows, zid, keys, abort_flag);
  else
    hipLaunchKernelGGL((pair_dist_kernel<KIND, STREAM, false>), dim3(grid), dim3(64), 0, s, queries, q0, d, X, woff,
                       qn, list, lstride, loff, nrows, zid, keys, abort_flag);
}

// Exclusive scans in two passes over blocks of 1024 entries: scan_local_kernel scans every block and
// records its total, scan_fix_kernel adds the totals of the blocks before it (there are at most
// kScanBlocks blocks: QUERY_CHUNK <= 65536).
constexpr int kScanBlocks = 64;

// Block-wide exclusive scan of NA values per thread (NA independent scans share the barriers).
template <int NA>
__device__ __forceinline__ void block_excl_scan(unsigned long long (&v)[NA], unsigned long long (*wsum)[16],
                                                unsigned long long (&total)[NA]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned long long x[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    x[a] = v[a];
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const unsigned long long y = __shfl_up(x[a], o, 64);
      if (lane >= o) x[a] += y;
    }
    if (lane == 63) wsum[a][wave] = x[a];
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    unsigned long long before = 0, tot = 0;
    for (int w2 = 0; w2 < 16; ++w2) {
      const unsigned long long t = wsum[a][w2];
      if (w2 < wave) before += t;
      tot += t;
    }
    total[a] = tot;
    v[a] = before + x[a] - v[a];
  }
}

// In-place local scans of up to three u64 arrays (null = absent); block totals to bsum[a][block].
// With `n32` given, array 0 is first filled from it as ceil(n32[i] / 64), 0 where mode[i] == skip_mode.
__global__ __launch_bounds__(1024) void scan_local_kernel(uint64_t* __restrict__ v0, uint64_t* __restrict__ v1,
                                                          uint64_t* __restrict__ v2, uint64_t n,
                                                          const uint32_t* __restrict__ n32,
                                                          const uint32_t* __restrict__ mode, uint32_t skip_mode,
                                                          unsigned long long* __restrict__ bsum,
    const uint32_t* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  __shared__ unsigned long long wsum[3][16];
  uint64_t* const vs[3] = {v0, v1, v2};
  const uint64_t i = (uint64_t)blockIdx.x * 1024 + threadIdx.x;
  unsigned long long x[3], tot[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) x[a] = (vs[a] && i < n) ? vs[a][i] : 0ull;
  if (n32) x[0] = (i < n && !(mode && mode[i] == skip_mode)) ? (n32[i] + 63u) / 64u : 0ull;
  block_excl_scan<3>(x, wsum, tot);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (vs[a] && i < n) vs[a][i] = x[a];
    if (threadIdx.x == 0) bsum[a * kScanBlocks + blockIdx.x] = tot[a];
  }
}

__global__ __launch_bounds__(1024) void scan_fix_kernel(uint64_t* __restrict__ v0, uint64_t* __restrict__ v1,
                                                        uint64_t* __restrict__ v2, uint64_t n,
                                                        const unsigned long long* __restrict__ bsum,
                                                        uint64_t cap0, uint64_t cap1, uint32_t* __restrict__ cap_abort,
    const uint32_t* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  uint64_t* const vs[3] = {v0, v1, v2};
  const uint64_t i = (uint64_t)blockIdx.x * 1024 + threadIdx.x;
  const uint32_t nb = (uint32_t)((n + 1023) / 1024);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (!vs[a]) continue;
    unsigned long long before = 0, all = 0;
    for (uint32_t b = 0; b < nb; ++b) {
      const unsigned long long t = bsum[a * kScanBlocks + b];
      if (b < blockIdx.x) before += t;
      all += t;
    }
    if (i < n) vs[a][i] += before;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      vs[a][n] = all;
      // graph replay: arrays 0 / 1 are the refine-list and ADC sizes, captured with capacities cap0 / cap1
      if (cap_abort && ((a == 0 && all > cap0) || (a == 1 && all > cap1))) *cap_abort = 1u;
    }
  }
}

// Sort keys[0..n) (global scratch) ascending into `lds` (capacity `cap` keys, a power of two) when n <= cap,
// or select the k smallest and sort those (k <= cap).  Returns how many sorted keys are available (min(n,k)).
template <class Src>
__device__ uint32_t select_sorted(Src g, uint32_t n, uint32_t k, uint64_t* lds, uint32_t cap, uint32_t* hist,
                                  uint64_t* sh) {
  uint32_t m = n < k ? n : k;
  if (n <= cap) {
    for (uint32_t i = threadIdx.x; i < n; i += kWG) lds[i] = g(i);
    __syncthreads();
    dev::wg_sort_lds(lds, n);
    return m;
  }
  if (m > cap) m = cap;  // callers size cap >= k
  dev::wg_radix_select(g, n, m, lds, hist, sh);
  dev::wg_sort_lds(lds, m);
  return m;
}

// LDS sort capacity for a top-k: at least 1024 keys and at least k, a power of two.
inline uint32_t sort_cap(uint64_t k) {
  uint32_t c = 1024;
  while (c < k) c <<= 1;
  return c;
}

// Number of leading sorted keys whose value is <= 1e6 (the keys ascend by value).
__device__ __forceinline__ uint32_t count_le_pad(const uint64_t* sorted, uint32_t m) {
  uint32_t a = 0, b = m;
  while (a < b) {
    const uint32_t mid = (a + b) >> 1;
    if (dev::key_value(sorted[mid]) <= kExactPad) a = mid + 1; else b = mid;
  }
  return a;
}

__device__ __forceinline__ uint32_t kept_with_pads(const uint64_t* sorted, uint32_t m, uint32_t k,
                                                   uint64_t pads) {
  const uint32_t n_le = count_le_pad(sorted, m);
  uint32_t keep = k < n_le ? k : n_le;
  if (k > n_le + pads) keep = n_le + (uint32_t)min<uint64_t>(k - n_le - pads, m - n_le);
  return keep;
}

// A5 select: top kc = min(n_list, group max nnz) of the exact cluster distances, sorted
// (Super-Grid-IVF_PQ.cu:1027-1042, refine.cu:78-119 with is_sort = true).  chosen_c holds positions in
// the query's non-zero cluster list.
__global__ __launch_bounds__(256) void coarse_pick_kernel(uint32_t C, uint32_t n_list, uint64_t G, uint64_t q0,
                                                          uint32_t scap,
                                                          const uint32_t* __restrict__ nnz_c,
                                                          const unsigned long long* __restrict__ gstat,
                                                          uint64_t* __restrict__ keys, uint32_t* __restrict__ chosen_c,
                                                          uint32_t* __restrict__ n_chosen_c) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint64_t* sorted = reinterpret_cast<uint64_t*>(smem);
  __shared__ __align__(16) uint32_t hist[2048];
  __shared__ uint64_t sh[8];
  const uint64_t qi = blockIdx.x;
  const uint32_t n = nnz_c[qi];
  const uint64_t gmax = gstat[((q0 + qi) / G) * 8 + 0];
  const uint32_t kc = (uint32_t)min<uint64_t>(n_list, gmax);
  const uint32_t m = select_sorted(dev::KeyArray{keys + qi * C}, n, kc, sorted, scap, hist, sh);
  __syncthreads();
  const uint32_t keep = kept_with_pads(sorted, m, kc, (uint64_t)C - n);
  for (uint32_t j = threadIdx.x; j < keep; j += kWG) chosen_c[qi * n_list + j] = dev::key_tie(sorted[j]);
  if (threadIdx.x == 0) n_chosen_c[qi] = keep;
}

// A6 part 1: the chosen clusters (in their sorted order) expanded into their non-zero sub-clusters in
// table order (build_offset_table, mat_operators.cuh:692-794; collect_segments_data +
// filter_none_zero_data, Super-Grid-IVF_PQ.cu:866-883), with their in-box counts.  The group maxima
// of the non-zero count and of the table width (all sub-clusters of the chosen clusters) feed A6's k
// and pad columns.
__global__ __launch_bounds__(256) void sub_expand_kernel(uint32_t CS, uint32_t C, const uint32_t* __restrict__ sub_off,
                                                         uint32_t n_list, uint64_t G, uint64_t q0,
                                                         const uint32_t* __restrict__ list_c,
                                                         const uint32_t* __restrict__ cl_beg,
                                                         const uint32_t* __restrict__ sub_list,
                                                         const uint32_t* __restrict__ sub_cnt,
                                                         const uint32_t* __restrict__ chosen_c,
                                                         const uint32_t* __restrict__ n_chosen_c, uint32_t cap,
                                                         uint32_t* __restrict__ list_s, uint32_t* __restrict__ list_scnt,
                                                         uint32_t* __restrict__ nnz_s,
                                                         unsigned long long* __restrict__ gstat) {
  __shared__ uint32_t tmp[8];
  const uint64_t qi = blockIdx.x;
  const uint32_t nc = n_chosen_c[qi];
  uint32_t base = 0, row_len = 0;
  for (uint32_t b = 0; b < nc; b += kWG) {
    const uint32_t i = b + threadIdx.x;
    uint32_t beg = 0, len = 0, tl = 0;
    if (i < nc) {
      const uint32_t j = chosen_c[qi * n_list + i];
      const uint32_t c = list_c[qi * C + j];
      beg = cl_beg[qi * (C + 1) + j];
      len = cl_beg[qi * (C + 1) + j + 1] - beg;
      tl = sub_off[c + 1] - sub_off[c];
    }
    uint32_t total, ttot;
    const uint32_t pos = dev::wg_exclusive_scan(len, tmp, total);
    (void)dev::wg_exclusive_scan(tl, tmp, ttot);
    for (uint32_t e = 0; e < len; ++e) {
      list_s[qi * cap + base + pos + e] = sub_list[qi * CS + beg + e];
      list_scnt[qi * cap + base + pos + e] = sub_cnt[qi * CS + beg + e];
    }
    base += total;
    row_len += ttot;
  }
  if (threadIdx.x == 0) {
    nnz_s[qi] = base;
    const uint64_t g = (q0 + qi) / G;
    atomicMax(&gstat[g * 8 + 1], (unsigned long long)base);
    atomicMax(&gstat[g * 8 + 2], (unsigned long long)row_len);
  }
}

// A6 part 2 select: top ks = min(sub_lists, group max) of the exact sub-center distances, with the
// group's table width supplying the 1e6 pad columns (Super-Grid-IVF_PQ.cu:885-896).  Output sorted
// ascending (the reference's order here is RAFT-defined) as positions in list_s.  With a per-query
// budget array the query keeps its own min(budget, sub_lists) nearest and no pads.
__global__ __launch_bounds__(256) void sub_pick_kernel(uint32_t sub_lists, uint64_t G, uint64_t q0, uint32_t cap,
                                                       uint32_t scap,
                                                       const uint32_t* __restrict__ nnz_s,
                                                       const unsigned long long* __restrict__ gstat,
                                                       const uint32_t* __restrict__ sub_budget,
                                                       uint64_t* __restrict__ keys, uint32_t* __restrict__ chosen_s,
                                                       uint32_t* __restrict__ n_chosen_s) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint64_t* sorted = reinterpret_cast<uint64_t*>(smem);
  __shared__ __align__(16) uint32_t hist[2048];
  __shared__ uint64_t sh[8];
  const uint64_t qi = blockIdx.x;
  const uint32_t n = nnz_s[qi];
  const uint64_t g = (q0 + qi) / G;
  uint32_t ks = (uint32_t)min<uint64_t>(sub_lists, gstat[g * 8 + 1]);
  uint64_t table_len = gstat[g * 8 + 2];
  if (sub_budget) {
    const uint32_t b = sub_budget[qi];
    ks = b == 0 || b > sub_lists ? sub_lists : b;
    table_len = n;
  }
  const uint32_t m = select_sorted(dev::KeyArray{keys + qi * cap}, n, ks, sorted, scap, hist, sh);
  __syncthreads();
  const uint32_t keep = kept_with_pads(sorted, m, ks, table_len - n);
  for (uint32_t j = threadIdx.x; j < keep; j += kWG) chosen_s[qi * sub_lists + j] = dev::key_tie(sorted[j]);
  if (threadIdx.x == 0) n_chosen_s[qi] = keep;
}

// ------------------------------------------------------------------------------------------------
// A8/A9: per chosen sub-cluster, the Morton range [lower_bound(zlo), lower_bound(zhi+1)) inside its
// z-pool segment (zorder.cu:656-685).  One workgroup per query, one thread per chosen sub-cluster.
// Also sums C_q (candidates) and the in-box count of the chosen sub-clusters (A10 inputs), and
// records the chosen sub-cluster ids (UINT32_MAX past the end).
// ranges_kernel's in-box branch (mode 3 queries: only the chosen sub-cluster ids and the in-box total) with
// one wave per query.
__global__ __launch_bounds__(256) void ranges_inbox_wave_kernel(
    uint64_t qn, uint32_t cap, const uint32_t* __restrict__ list_s, const uint32_t* __restrict__ list_scnt,
    uint32_t sub_lists, const uint32_t* __restrict__ chosen_s, const uint32_t* __restrict__ n_chosen_s,
    uint32_t* __restrict__ chosen_sid, uint32_t* __restrict__ rpre, uint64_t* __restrict__ cand,
    uint64_t* __restrict__ real, const uint32_t* __restrict__ n_ib,
    const uint32_t* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  const int lane = threadIdx.x & 63;
  const uint64_t qi = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (qi >= qn || n_ib[qi] == kNone) return;
  const uint32_t n = n_chosen_s[qi];
  uint64_t r = 0;
  for (uint32_t j = lane; j < sub_lists; j += 64) {
    uint32_t sid = kNone;
    if (j < n) {
      const uint32_t idx = chosen_s[qi * sub_lists + j];
      sid = list_s[qi * cap + idx];
      r += list_scnt[qi * cap + idx];
    }
    chosen_sid[qi * sub_lists + j] = sid;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) r += dev::shfl_xor_u64(r, o);
  if (lane == 0) {
    cand[qi] = 0;
    real[qi] = r;
    rpre[qi * (sub_lists + 1)] = 0;
  }
}

__global__ __launch_bounds__(256) void ranges_kernel(
    const uint64_t* __restrict__ zcode, const uint32_t* __restrict__ seg_off, const uint32_t* __restrict__ cells,
    uint64_t N, int n_valid, const uint32_t* __restrict__ box_lo, const uint32_t* __restrict__ box_w, uint64_t qn,
    const uint64_t* __restrict__ zlo_cnt, uint32_t cap,
    const uint32_t* __restrict__ list_s, const uint32_t* __restrict__ list_scnt, uint32_t sub_lists, uint64_t G,
    uint64_t q0, const uint32_t* __restrict__ chosen_s, const uint32_t* __restrict__ n_chosen_s,
    const uint64_t* __restrict__ zlo, const uint64_t* __restrict__ zhi, uint32_t* __restrict__ chosen_sid,
    uint32_t* __restrict__ rstart, uint32_t* __restrict__ rpre, uint64_t* __restrict__ cand,
    uint64_t* __restrict__ real, unsigned long long* __restrict__ gstat, const uint32_t* __restrict__ n_ib,
    int inbox_mode, int split,
    const uint32_t* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  __shared__ uint32_t tmp[8];
  __shared__ uint32_t base_sh;
  __shared__ unsigned long long real_sh;
  __shared__ uint32_t cs_sh[kWG], cp_sh[kWG];
  const uint64_t qi = blockIdx.x;
  const uint32_t n = n_chosen_s[qi];
  if (inbox_mode == 2 && n_ib[qi] != kNone) return;  // done by ranges_inbox_wave_kernel
  if (inbox_mode && n_ib[qi] != kNone) {
    // refined from its in-box list (mode 3): only the chosen ids and the in-box total are needed
    unsigned long long r = 0;
    for (uint32_t j = threadIdx.x; j < sub_lists; j += kWG) {
      uint32_t sid = kNone;
      if (j < n) {
        const uint32_t idx = chosen_s[qi * sub_lists + j];
        sid = list_s[qi * cap + idx];
        r += list_scnt[qi * cap + idx];
      }
      chosen_sid[qi * sub_lists + j] = sid;
    }
    if (threadIdx.x == 0) real_sh = 0;
    __syncthreads();
    atomicAdd(&real_sh, r);
    __syncthreads();
    if (threadIdx.x == 0) {
      cand[qi] = 0;
      real[qi] = real_sh;
      rpre[qi * (sub_lists + 1)] = 0;
    }
    return;
  }
  const uint64_t a = zlo[qi], b = zhi[qi] + 1;
  if (threadIdx.x == 0) {
    base_sh = 0;
    real_sh = 0;
  }
  __syncthreads();
  const bool dense_q = n_ib[qi] == kNone;
  uint32_t blo[kMaxDims], bw[kMaxDims];
  if (dense_q)
    for (int j2 = 0; j2 < n_valid; ++j2) {
      blo[j2] = box_lo[(uint64_t)j2 * qn + qi];
      bw[j2] = box_w[(uint64_t)j2 * qn + qi];
    }
  for (uint32_t j0 = 0; j0 < sub_lists; j0 += kWG) {
    const uint32_t j = j0 + threadIdx.x;
    uint32_t cnt = 0, start = 0, sid = kNone, cstart = 0, clen = 0;
    if (j < n) {
      const uint32_t idx = chosen_s[qi * sub_lists + j];
      sid = list_s[qi * cap + idx];
      const uint32_t p0 = seg_off[sid], p1 = seg_off[sid + 1];
      // lower_bound(zlo) and lower_bound(zhi + 1) side by side, so their loads overlap
      uint32_t lo = p0, hi = p1, lo2 = p0, hi2 = p1;
      while (lo < hi || lo2 < hi2) {
        const uint32_t m1 = (lo + hi) >> 1, m2 = (lo2 + hi2) >> 1;
        const uint64_t z1 = lo < hi ? zcode[m1] : 0, z2 = lo2 < hi2 ? zcode[m2] : 0;
        if (lo < hi) {
          if (z1 < a) lo = m1 + 1; else hi = m1;
        }
        if (lo2 < hi2) {
          if (z2 < b) lo2 = m2 + 1; else hi2 = m2;
        }
      }
      start = lo;
      cnt = lo2 > lo ? lo2 - lo : 0u;  // an empty code range (zlo > zhi) holds no candidate
      if (dense_q) {
        // flagged by the dense pass: the box is counted below over the count range [z(max(lo, 0)), z(hi)]
        uint32_t c0 = lo;
        if (zlo_cnt[qi] != a) {
          uint32_t x = p0, y = p1;
          while (x < y) {
            const uint32_t mid = (x + y) >> 1;
            if (zcode[mid] < zlo_cnt[qi]) x = mid + 1; else y = mid;
          }
          c0 = x;
        }
        cstart = c0;
        clen = lo2 > c0 ? lo2 - c0 : 0u;
      } else {
        atomicAdd(&real_sh, (unsigned long long)list_scnt[qi * cap + idx]);
      }
    }
    if (dense_q) {
      // the chunk's count ranges flattened over the workgroup
      uint32_t ctot;
      const uint32_t cpre = dev::wg_exclusive_scan(clen, tmp, ctot);
      cs_sh[threadIdx.x] = cstart;
      cp_sh[threadIdx.x] = cpre;
      __syncthreads();
      // a thread's positions t ascend: one search for the owner of the first, then step forward; four
      // positions per step, each dim's four cell loads in flight together
      uint32_t mine = 0;
      uint32_t x = 0;  // last owner k with cp_sh[k] <= t
      {
        uint32_t y = kWG;
        while (y - x > 1) {
          const uint32_t mid = (x + y) >> 1;
          if (cp_sh[mid] <= threadIdx.x) x = mid; else y = mid;
        }
      }
      for (uint32_t t0 = threadIdx.x; t0 < ctot; t0 += 4 * kWG) {
        uint32_t p[4];
        bool in[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t t = t0 + (uint32_t)u * kWG;
          in[u] = t < ctot;
          p[u] = 0;
          if (in[u]) {
            while (x + 1 < (uint32_t)kWG && cp_sh[x + 1] <= t) ++x;
            p[u] = cs_sh[x] + (t - cp_sh[x]);
          }
        }
        for (int j2 = 0; j2 < n_valid; ++j2) {
          const uint32_t* cb = cells + (uint64_t)j2 * N;
          uint32_t c[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) c[u] = in[u] ? cb[p[u]] : 0u;
#pragma unroll
          for (int u = 0; u < 4; ++u) in[u] = in[u] && c[u] - blo[j2] <= bw[j2];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) mine += in[u] ? 1u : 0u;
      }
      atomicAdd(&real_sh, (unsigned long long)mine);
      __syncthreads();
    }
    uint32_t total;
    uint32_t pos = dev::wg_exclusive_scan(cnt, tmp, total);
    if (j < sub_lists) {
      // split mode: the range table is written by split_ranges_kernel
      if (!split) {
        rstart[qi * sub_lists + j] = start;
        rpre[qi * (sub_lists + 1) + j] = base_sh + pos;
      }
      chosen_sid[qi * sub_lists + j] = sid;
    }
    __syncthreads();
    if (threadIdx.x == 0) base_sh += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const uint64_t g = (q0 + qi) / G;
    real[qi] = real_sh;
    atomicAdd(&gstat[g * 8 + 4], real_sh);
    if (!split) {
      rpre[qi * (sub_lists + 1) + sub_lists] = base_sh;
      cand[qi] = base_sh;
      atomicAdd(&gstat[g * 8 + 3], (unsigned long long)base_sh);
      atomicMax(&gstat[g * 8 + 5], (unsigned long long)base_sh);
    }
  }
}

// ------------------------------------------------------------------------------------------------
// A8 in Morton split mode (is_split = 1, zorder.cu:474-654): each chosen sub-cluster's Morton range is cut into
// N_SPLIT near-equal shares (get_query_end_points_kernel, :591-647); each share starts at its first point and
// jumps at most three times to the next in-box code (BIGMIN) until it reaches an in-box point (the share then
// starts there) or no in-box code is left before its last code (the share is dropped: the reference compares
// `next >= last code`).  The candidate sub-cluster is the query's own (the reference's kernel reads candidates[c],
// H23).  The query's range table holds sub_lists * N_SPLIT pieces, chosen sub-cluster major; n_rng = pieces in use.
// BIGMIN here is a bit scan over the interleaved code (Tropf and Herzog): the smallest code above `cur` whose
// point lies in the box [zmin, zmax]; the reference's step computes the same code for every cur below the largest.
__device__ __forceinline__ bool bigmin_next(uint64_t cur, uint64_t zmin, uint64_t zmax, int nb, int a,
                                            const uint64_t* dmask, uint64_t& out) {
  if (cur == ~0ull || (nb < 64 && cur + 1 >= (1ull << nb))) return false;  // no code above the largest one
  const uint64_t z = cur + 1;
  uint64_t best = 0;
  bool have = false;
  for (int b = nb - 1; b >= 0; --b) {
    const uint64_t bit = 1ull << b;
    const bool zb = (z >> b) & 1ull, lb = (zmin >> b) & 1ull, hb = (zmax >> b) & 1ull;
    const uint64_t m = dmask[(nb - 1 - b) % a] & ((bit << 1) - 1ull);  // this dim's bits at or below b
    if (!zb && !lb && hb) {
      best = (zmin & ~m) | bit;            // lowest code of the box's upper half in this dim
      have = true;
      zmax = (zmax & ~m) | (m & ~bit);     // the box's lower half in this dim
    } else if (!zb && lb && hb) {
      out = zmin;
      return true;
    } else if (zb && !lb && !hb) {
      out = best;
      return have;
    } else if (zb && !lb && hb) {
      zmin = (zmin & ~m) | bit;            // the box's upper half in this dim
    }
  }
  out = z;  // z itself lies in the box
  return true;
}

__global__ __launch_bounds__(256) void split_ranges_kernel(
    const uint64_t* __restrict__ zcode, const uint32_t* __restrict__ seg_off, const uint32_t* __restrict__ cells,
    uint64_t N, int n_valid, int max_bits, const uint32_t* __restrict__ box_lo, const uint32_t* __restrict__ box_w,
    uint64_t qn, const uint64_t* __restrict__ zlo, const uint64_t* __restrict__ zhi,
    const uint64_t* __restrict__ zlo_cnt, uint32_t sub_lists, uint32_t n_split,
    const uint32_t* __restrict__ chosen_sid, const uint32_t* __restrict__ n_chosen_s,
    const uint32_t* __restrict__ n_ib, int inbox_mode, uint32_t* __restrict__ rstart, uint32_t* __restrict__ rpre,
    uint32_t* __restrict__ n_rng, uint64_t* __restrict__ cand, unsigned long long* __restrict__ gstat, uint64_t G,
    uint64_t q0, uint32_t* __restrict__ err, const uint32_t* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  __shared__ uint32_t tmp[8];
  const uint64_t qi = blockIdx.x;
  if (inbox_mode && n_ib[qi] != kNone) {  // refined from its in-box list (mode 3)
    if (threadIdx.x == 0) n_rng[qi] = 0;
    return;
  }
  const uint32_t R = sub_lists * n_split;
  const uint32_t items = n_chosen_s[qi] * n_split;
  const uint64_t za = zlo[qi], zb = zhi[qi];
  // a lower corner beyond the label range wrapped: the reference's step then compares mismatched boxes
  if (threadIdx.x == 0 && za != zlo_cnt[qi]) atomicOr(err, 2u);
  const int nb = n_valid * max_bits;
  uint64_t dmask[kMaxDims];
  for (int j = 0; j < n_valid; ++j) {
    uint64_t m = 0;
    for (int p = 0; p < nb; ++p)
      if ((nb - 1 - p) % n_valid == j) m |= 1ull << p;
    dmask[j] = m;
  }
  uint32_t blo[kMaxDims], bw[kMaxDims];
  for (int j = 0; j < n_valid; ++j) {
    blo[j] = box_lo[(uint64_t)j * qn + qi];
    bw[j] = box_w[(uint64_t)j * qn + qi];
  }
  uint32_t base = 0;
  for (uint32_t t0 = 0; t0 < R; t0 += kWG) {
    const uint32_t t = t0 + threadIdx.x;
    uint32_t start = 0, len = 0;
    if (t < items) {
      const uint32_t j = t / n_split, sp = t - j * n_split;
      const uint32_t sid = chosen_sid[qi * sub_lists + j];
      const uint32_t p0 = seg_off[sid], p1 = seg_off[sid + 1];
      uint32_t lo = p0, hi = p1, lo2 = p0, hi2 = p1;
      while (lo < hi || lo2 < hi2) {
        const uint32_t m1 = (lo + hi) >> 1, m2 = (lo2 + hi2) >> 1;
        const uint64_t z1 = lo < hi ? zcode[m1] : 0, z2 = lo2 < hi2 ? zcode[m2] : 0;
        if (lo < hi) {
          if (z1 < za) lo = m1 + 1; else hi = m1;
        }
        if (lo2 < hi2) {
          if (z2 < zb + 1) lo2 = m2 + 1; else hi2 = m2;
        }
      }
      const uint32_t L = lo2 > lo ? lo2 - lo : 0u;
      const uint32_t ns = L < n_split ? L : n_split;
      if (sp < ns) {
        const uint32_t qt = L / ns, r = L - qt * ns;
        uint32_t a = lo + sp * qt + (sp < r ? sp : r);
        const uint32_t b = a + qt + (sp < r ? 1u : 0u);
        const uint64_t zlast = zcode[b - 1];
        uint64_t cur = zcode[a];
        bool valid = true;
        for (int jumps = 0; jumps < 3;) {
          uint32_t x = a, y = b;  // lower_bound(cur) in [a, b)
          while (x < y) {
            const uint32_t mid = (x + y) >> 1;
            if (zcode[mid] < cur) x = mid + 1; else y = mid;
          }
          if (x >= b) break;
          a = x;
          bool in = true;
          for (int jd = 0; jd < n_valid; ++jd) in = in && (cells[(uint64_t)jd * N + x] - blo[jd] <= bw[jd]);
          ++jumps;
          if (in) break;
          uint64_t nz = 0;
          if (!bigmin_next(zcode[x], za, zb, nb, n_valid, dmask, nz) || nz >= zlast) {
            valid = false;
            break;
          }
          cur = nz;
        }
        start = a;
        len = valid ? b - a : 0u;
      }
    }
    uint32_t total;
    const uint32_t pos = dev::wg_exclusive_scan(len, tmp, total);
    if (t < R) {
      rstart[qi * R + t] = start;
      rpre[qi * (R + 1) + t] = base + pos;
    }
    base += total;
  }
  if (threadIdx.x == 0) {
    rpre[qi * (R + 1) + R] = base;
    cand[qi] = base;
    n_rng[qi] = items;
    const uint64_t g = (q0 + qi) / G;
    atomicAdd(&gstat[g * 8 + 3], (unsigned long long)base);
    atomicMax(&gstat[g * 8 + 5], (unsigned long long)base);
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
// A11 selection over many workgroups.  A mode-1 query with at least ADC_SPLIT ADC candidates gets a slot
// (kSelSlots per chunk: every query of the largest chunk).  The scan records the range [mn, mx] of its distance bits; bins are those of
// dev::dist_bin_shift (monotone in the distance).  Workgroups of kSelTile candidates add their LDS histograms into the slot's global histogram,
// then split the candidates into keys below the k'-th key's bin (final) and keys in that bin (scratch); the
// per-query selection finally ranks the keys of that one bin.
constexpr uint32_t kSelTile = 16384;
constexpr uint32_t kSelSlots = 65536;
constexpr uint32_t kSelBins = 2048;

__global__ void plan_kernel(uint64_t qn, uint64_t q0, uint64_t G, int l, uint64_t exps0, uint64_t topk,
                            const uint64_t* __restrict__ cand, const unsigned long long* __restrict__ gstat,
                            uint32_t* __restrict__ mode, uint64_t* __restrict__ kprime,
                            uint64_t* __restrict__ lsize, uint64_t* __restrict__ asize,
                            uint64_t* __restrict__ chunks, uint64_t* __restrict__ pad32,
                            uint64_t* __restrict__ padmax, const uint32_t* __restrict__ n_ib, int inbox_mode,
                            uint32_t adc_chunk, uint32_t* __restrict__ sslot, uint32_t* __restrict__ smm,
                            uint32_t* __restrict__ scnt, uint32_t* __restrict__ slot_ctr, uint32_t slot_cap,
                            uint64_t split_min,
    const uint32_t* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  uint64_t qi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (qi >= qn) return;
  sslot[qi] = kNone;
  smm[2 * qi] = 0xFFFFFFFFu;  // distance-bit range of the ADC scan, filled by the scan
  smm[2 * qi + 1] = 0u;
  scnt[4 * qi] = 0u;
  scnt[4 * qi + 1] = 0u;
  scnt[4 * qi + 2] = 0u;
  if (inbox_mode && n_ib[qi] != kNone) {
    // mode 3 (mixed precision, walked query): exact re-rank of the chosen sub-clusters' in-box points that
    // pass the label predicate
    mode[qi] = 3;
    kprime[qi] = 0;
    lsize[qi] = n_ib[qi];
    asize[qi] = 0;
    chunks[qi] = 0;
    pad32[qi] = 0;
    padmax[qi] = 0;
    return;
  }
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
    chunks[qi] = cdiv(Cq, adc_chunk);
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
  chunks[qi] = adc ? cdiv(Cq, adc_chunk) : 0;
  if (adc && split_min && Cq >= split_min) {
    const uint32_t sl = atomicAdd(slot_ctr, 1u);
    if (sl < slot_cap) sslot[qi] = sl;  // beyond the slots: the one-workgroup selection
  }
  pad32[qi] = p32;
  padmax[qi] = pmax;  // l > 1: set after the post-filter (n_candies - survivors)
}

// Position in the z-pool of candidate column `col` of a query row (rstart/rpre of that query, in LDS).
__device__ __forceinline__ uint32_t col_to_pos(const uint32_t* rs, const uint32_t* rp, uint32_t n,
                                               uint32_t col) {
  uint32_t lo = 0, hi = n;  // last j with rp[j] <= col
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (rp[mid] <= col) lo = mid; else hi = mid;
  }
  return rs[lo] + (col - rp[lo]);
}

// Label words (IndexState::labw): the label predicate on integer labels as bounds on the word's fields.
// For an integer v = minv + f and f32 bounds a, b: a <= v <= b exactly when ceil(a) - minv <= f <=
// floor(b) - minv; a NaN bound fails no comparison, so it does not bound.  label_box turns a query's 2 l
// bounds into field values LO (lower bounds) and SPAN (upper - lower), and all fields are tested at once with
// the zero guard bit above each field (G): X = (w | G) - LO keeps field j's guard set exactly when f_j >= lo_j
// (no borrow crosses a guard), its field then holding f_j - lo_j; Z = (SPAN | G) - (X & ~G) keeps it set
// exactly when f_j - lo_j <= span_j.  A dim no value can pass sets `none`.
struct LabelWords {
  const uint64_t* w;            // nullptr: read the f32 label rows
  const LabelWordLayout* lay;   // device copy of the layout (read once per query)
  uint64_t guard;               // lay->guard
};

struct LabelBox {
  uint64_t lo, span;
  uint32_t none;
};

__device__ inline LabelBox label_box(const LabelWords& P, int l, const float* rg) {
  uint64_t lo_w = 0, span_w = 0;
  bool empty = false;
#pragma unroll
  for (int j = 0; j < kMaxDims; ++j) {
    if (j >= l) continue;
    const double m = (double)P.lay->minv[j];
    const double a = rg[2 * j], b = rg[2 * j + 1];
    double lo = 0.0, hi = (double)P.lay->mask[j];
    if (a == a) lo = fmax(lo, ceil(a) - m);
    if (b == b) hi = fmin(hi, floor(b) - m);
    if (lo > hi) {
      empty = true;
    } else {
      lo_w |= (uint64_t)(uint32_t)lo << P.lay->off[j];
      span_w |= (uint64_t)(uint32_t)(hi - lo) << P.lay->off[j];
    }
  }
  // every caller's query is uniform over the wave: the bounds live in scalar registers
  // (readfirstlane returns a signed int: each half goes through uint32_t before it is widened)
  auto uniform_u64 = [](uint64_t v) {
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    return ((uint64_t)hi << 32) | (uint64_t)lo;
  };
  LabelBox B;
  B.lo = uniform_u64(lo_w);
  B.span = uniform_u64(span_w);
  B.none = (uint32_t)__builtin_amdgcn_readfirstlane(empty ? 1 : 0);
  return B;
}

__device__ __forceinline__ bool label_word_pass(const LabelWords& P, const LabelBox& B, uint64_t w) {
  const uint64_t G = P.guard;
  const uint64_t x = (w | G) - B.lo;
  const uint64_t z = (B.span | G) - (x & ~G);
  return !B.none && (x & z & G) == G;
}

// The exact label predicate lo <= label <= hi in every dim (filter.cu:30-64).  Up to four dims are loaded
// together (one memory round trip); more dims are tested one by one.
__device__ __forceinline__ bool label_pass(const float* __restrict__ lab, int l, uint32_t pos,
                                           const float* rg) {
  const float* lp = lab + (uint64_t)pos * l;
  if (l <= 4) {
    float v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = i < l ? lp[i] : 0.0f;
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (i < l && (v[i] < rg[2 * i] || v[i] > rg[2 * i + 1])) ok = false;
    return ok;
  }
  for (int i = 0; i < l; ++i) {
    const float v = lp[i];
    if (v < rg[2 * i] || v > rg[2 * i + 1]) return false;
  }
  return true;
}

// Mode 0 / 2: copy every candidate (l == 1) or every candidate passing the exact label predicate
// (l > 1 and the query_prefilter path) into the refine list, in candidate-column order
// (filter.cu:146-164).  One workgroup per query; the query's range table is staged in LDS.
__device__ __forceinline__ bool inbox_keeps(uint64_t qi, int l, int integral, const uint32_t* __restrict__ mode,
                                            const uint32_t* __restrict__ n_ib, const uint64_t* __restrict__ zlo,
                                            const uint64_t* __restrict__ zlo_cnt) {
  // integral < 0 (Morton split mode): the keep-all candidates are the split pieces, so they are never replaced
  return mode[qi] == 3 ||
         (mode[qi] == 0 && integral >= 0 && n_ib[qi] != kNone && zlo[qi] == zlo_cnt[qi] && (l == 1 || integral));
}

__global__ __launch_bounds__(256) void keep_all_kernel(
    uint64_t qn, int l, int integral, uint32_t sub_lists, const uint32_t* __restrict__ mode,
    const uint32_t* __restrict__ n_ib, const uint64_t* __restrict__ zlo, const uint64_t* __restrict__ zlo_cnt,
    const uint32_t* __restrict__ n_chosen_s, const uint32_t* __restrict__ rstart,
    const uint32_t* __restrict__ rpre, const float* __restrict__ lab, LabelWords lw,
    const float* __restrict__ ranges, const uint64_t* __restrict__ loff, uint32_t* __restrict__ L,
    uint32_t* __restrict__ nL, const uint32_t* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  __shared__ uint32_t tmp[8];
  extern __shared__ uint32_t rsp[];  // rs[sub_lists], rp[sub_lists + 1]
  uint32_t* rs = rsp;
  uint32_t* rp = rsp + sub_lists;
  const uint64_t qi = blockIdx.x;
  if (mode[qi] == 1 || inbox_keeps(qi, l, integral, mode, n_ib, zlo, zlo_cnt)) return;
  const bool filt = (l > 1) || mode[qi] == 2;
  const uint32_t nr = n_chosen_s[qi];
  const uint32_t Cq = rpre[qi * (sub_lists + 1) + nr];
  uint32_t* out = L + loff[qi];
  for (uint32_t j = threadIdx.x; j <= nr; j += kWG) {
    rp[j] = rpre[qi * (sub_lists + 1) + j];
    if (j < nr) rs[j] = rstart[qi * sub_lists + j];
  }
  __syncthreads();
  if (!filt) {
    for (uint32_t j = 0; j < nr; ++j) {
      const uint32_t p0 = rs[j], n = rp[j + 1] - rp[j];
      for (uint32_t e = threadIdx.x; e < n; e += kWG) out[rp[j] + e] = p0 + e;
    }
    if (threadIdx.x == 0) nL[qi] = Cq;
    return;
  }
  __shared__ float rg[2 * kMaxDims];
  if (threadIdx.x < 2 * (uint32_t)l) rg[threadIdx.x] = ranges[qi * 2 * l + threadIdx.x];
  __syncthreads();
  const LabelBox box = lw.w ? label_box(lw, l, rg) : LabelBox{};
  // each thread owns kKeepPer consecutive columns, so all their label loads are in flight together and
  // one scan per kWG * kKeepPer columns keeps the column order
  constexpr int kKeepPer = 8;
  uint32_t base = 0;
  for (uint32_t b = 0; b < Cq; b += kWG * kKeepPer) {
    const uint32_t c0 = b + threadIdx.x * kKeepPer;
    uint32_t pos[kKeepPer];
    uint32_t flags = 0, cnt = 0;
    if (c0 < Cq) {
      uint32_t lo = 0, hi = nr;  // last range j with rp[j] <= c0
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (rp[mid] <= c0) lo = mid; else hi = mid;
      }
      uint32_t j = lo;
#pragma unroll
      for (int i = 0; i < kKeepPer; ++i) {
        const uint32_t col = c0 + i;
        while (j + 1 < nr && rp[j + 1] <= col) ++j;
        pos[i] = rs[j] + (col - rp[j]);
      }
      // the label filter (label_pass) with every item's loads issued before the tests
#pragma unroll
      for (int i = 0; i < kKeepPer; ++i)
        if (c0 + i < Cq) flags |= 1u << i;
      if (lw.w) {
        uint64_t wv[kKeepPer];
#pragma unroll
        for (int i = 0; i < kKeepPer; ++i) wv[i] = (flags >> i) & 1u ? lw.w[pos[i]] : 0ull;
#pragma unroll
        for (int i = 0; i < kKeepPer; ++i)
          if (!label_word_pass(lw, box, wv[i])) flags &= ~(1u << i);
      } else if (l == 4) {
#pragma unroll
        for (int i = 0; i < kKeepPer; i += 4) {
          float4 lv[4];
#pragma unroll
          for (int u = 0; u < 4; ++u)
            lv[u] = (flags >> (i + u)) & 1u ? reinterpret_cast<const float4*>(lab)[pos[i + u]]
                                            : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const bool fail = lv[u].x < rg[0] || lv[u].x > rg[1] || lv[u].y < rg[2] || lv[u].y > rg[3] ||
                              lv[u].z < rg[4] || lv[u].z > rg[5] || lv[u].w < rg[6] || lv[u].w > rg[7];
            if (fail) flags &= ~(1u << (i + u));
          }
        }
      } else {
        for (int dm = 0; dm < l; ++dm) {
          float lv[kKeepPer];
#pragma unroll
          for (int i = 0; i < kKeepPer; ++i) lv[i] = (flags >> i) & 1u ? lab[(uint64_t)pos[i] * l + dm] : 0.f;
          const float lo_b = rg[2 * dm], hi_b = rg[2 * dm + 1];
#pragma unroll
          for (int i = 0; i < kKeepPer; ++i)
            if (lv[i] < lo_b || lv[i] > hi_b) flags &= ~(1u << i);
        }
      }
      cnt = (uint32_t)__popc(flags);
    }
    uint32_t total;
    uint32_t at = base + dev::wg_exclusive_scan(cnt, tmp, total);
#pragma unroll
    for (int i = 0; i < kKeepPer; ++i)
      if (flags & (1u << i)) out[at++] = pos[i];
    base += total;
  }
  if (threadIdx.x == 0) nL[qi] = base;
}

// Keep-all refine list from the query's in-box points (mode 0, a walked query whose Morton lower corner
// did not wrap): every label survivor among the Morton-range candidates lies in the cell box (labels are
// integers, or l == 1 where the range is the box), so the survivors are exactly the in-box points of the
// chosen sub-clusters that pass the label predicate (none is filtered when l == 1).  The list order differs
// from the candidate-column order, which the exact (distance, id) top-k does not depend on.  One
// workgroup per query; the chosen sub-clusters are a bit set in LDS.
__global__ __launch_bounds__(256) void inbox_keep_kernel(
    int l, int integral, uint32_t CS, uint32_t sub_lists, const uint32_t* __restrict__ mode,
    const uint32_t* __restrict__ n_ib, const uint64_t* __restrict__ zlo, const uint64_t* __restrict__ zlo_cnt,
    const uint32_t* __restrict__ n_chosen_s, const uint32_t* __restrict__ chosen_sid, uint32_t ibcap,
    const uint32_t* __restrict__ ib_pos, const uint32_t* __restrict__ ib_sub, const uint32_t* __restrict__ seg_off,
    uint32_t pk_shift, const float* __restrict__ lab, LabelWords lw, const float* __restrict__ ranges,
    const uint64_t* __restrict__ loff, uint32_t* __restrict__ L, uint32_t* __restrict__ nL,
    const uint32_t* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  __shared__ uint32_t tmp[8];
  __shared__ uint32_t bits[kSparseMaxCS / 32];
  const uint64_t qi = blockIdx.x;
  if (!inbox_keeps(qi, l, integral, mode, n_ib, zlo, zlo_cnt)) return;
  const uint32_t nw = (CS + 31) / 32;
  for (uint32_t i = threadIdx.x; i < nw; i += kWG) bits[i] = 0;
  __syncthreads();
  const uint32_t nc = n_chosen_s[qi];
  for (uint32_t j = threadIdx.x; j < nc; j += kWG) {
    const uint32_t sid = chosen_sid[qi * sub_lists + j];
    atomicOr(&bits[sid >> 5], 1u << (sid & 31));
  }
  __syncthreads();
  float rg[2 * kMaxDims];
  for (int i = 0; i < 2 * l; ++i) rg[i] = ranges[qi * 2 * l + i];
  const LabelBox box = lw.w ? label_box(lw, l, rg) : LabelBox{};
  const uint32_t n = n_ib[qi];
  const uint32_t md = mode[qi];
  uint32_t* out = L + loff[qi];
  uint32_t base = 0;
  for (uint32_t b = 0; b < n; b += kWG) {
    const uint32_t i = b + threadIdx.x;
    uint32_t pos = 0, flag = 0;
    if (i < n) {
      const uint32_t v = ib_pos[qi * ibcap + i];
      const uint32_t sid = ib_sub ? ib_sub[qi * ibcap + i] : v >> pk_shift;
      if ((bits[sid >> 5] >> (sid & 31)) & 1u) {
        pos = ib_sub ? v : seg_off[sid] + (v & ((1u << pk_shift) - 1u));
        flag = ((l == 1 && md != 3) ||
                (lw.w ? label_word_pass(lw, box, lw.w[pos]) : label_pass(lab, l, pos, rg))) ? 1u : 0u;
      }
    }
    uint32_t total;
    const uint32_t at = dev::wg_exclusive_scan(flag, tmp, total);
    if (flag) out[base + at] = pos;
    base += total;
  }
  if (threadIdx.x == 0) nL[qi] = base;
}

// inbox_keep_kernel with one wave per query (kKeepWaves per workgroup), each with its own chosen-sub-cluster bit
// set in LDS; the kept points are compacted by ballots in list order, so the output equals the workgroup
// kernel's.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

constexpr int kKeepWaves = 8;  // queries (waves) per workgroup: they share the staged seg_off

template <bool PACKED, bool WORDS>
__global__ __launch_bounds__(64 * kKeepWaves) void inbox_keep_wave_kernel(
    uint64_t qn, int l, int integral, uint32_t CS, uint32_t sub_lists, const uint32_t* __restrict__ mode,
    const uint32_t* __restrict__ n_ib, const uint64_t* __restrict__ zlo, const uint64_t* __restrict__ zlo_cnt,
    const uint32_t* __restrict__ n_chosen_s, const uint32_t* __restrict__ chosen_sid, uint32_t ibcap,
    const uint32_t* __restrict__ ib_pos, const uint32_t* __restrict__ ib_sub, const uint32_t* __restrict__ seg_off,
    uint32_t pk_shift, const float* __restrict__ lab, LabelWords lw, const float* __restrict__ ranges,
    const uint64_t* __restrict__ loff, uint32_t* __restrict__ L, uint32_t* __restrict__ nL,
    const uint32_t* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  __shared__ uint32_t bits_all[kKeepWaves][kSparseMaxCS / 32];
  extern __shared__ uint32_t seg_lds[];  // packed entries: seg_off[0 .. CS) for the position of a kept entry
  if (PACKED) {
    for (uint32_t i = threadIdx.x; i < CS; i += blockDim.x) seg_lds[i] = seg_off[i];
    __syncthreads();
  }
  const int lane = threadIdx.x & 63;
  const uint64_t qi = (uint64_t)blockIdx.x * kKeepWaves + (threadIdx.x >> 6);
  if (qi >= qn) return;
  if (!inbox_keeps(qi, l, integral, mode, n_ib, zlo, zlo_cnt)) return;
  uint32_t* bits = bits_all[threadIdx.x >> 6];
  const uint32_t nw = (CS + 31) / 32;
  for (uint32_t i = lane; i < nw; i += 64) bits[i] = 0;
  wave_lds_sync();
  const uint32_t nc = n_chosen_s[qi];
  for (uint32_t j = lane; j < nc; j += 64) {
    const uint32_t sid = chosen_sid[qi * sub_lists + j];
    atomicOr(&bits[sid >> 5], 1u << (sid & 31));
  }
  wave_lds_sync();
  float rg[2 * kMaxDims];
  for (int i = 0; i < 2 * l; ++i) rg[i] = ranges[qi * 2 * l + i];
  const LabelBox box = WORDS ? label_box(lw, l, rg) : LabelBox{};
  const uint32_t n = n_ib[qi];
  const uint32_t md = mode[qi];
  const bool filt = !(l == 1 && md != 3);
  uint32_t* out = L + loff[qi];
  uint32_t base = 0;
  // kKeepBatches batches of 64 list entries per pass: their list loads, then the label loads of the entries in
  // chosen sub-clusters, are in flight together; the kept entries are compacted batch by batch in list order.
  // The next pass's list entries are loaded before this pass's label loads, so a pass waits on one round trip.
  // Every load is issued unconditionally (indices clamped into the list / entry 0), so a batch's loads are in
  // flight together.
  constexpr int kKeepBatches = 8;
  uint32_t nxt_p[kKeepBatches], nxt_s[kKeepBatches];
  auto load_list = [&](uint32_t b0) {
#pragma unroll
    for (int u = 0; u < kKeepBatches; ++u) {
      const uint32_t i = b0 + 64u * u + (uint32_t)lane;
      const uint64_t at = qi * ibcap + (i < n ? i : n - 1);
      nxt_p[u] = ib_pos[at];
      if (!PACKED) nxt_s[u] = ib_sub[at];
    }
  };
  if (n > 0) load_list(0);
  for (uint32_t b0 = 0; b0 < n; b0 += 64u * kKeepBatches) {
    uint32_t sid[kKeepBatches], pos[kKeepBatches];
#pragma unroll
    for (int u = 0; u < kKeepBatches; ++u) {
      const uint32_t i = b0 + 64u * u + (uint32_t)lane;
      pos[u] = nxt_p[u];
      sid[u] = i >= n ? kNone : (PACKED ? pos[u] >> pk_shift : nxt_s[u]);
    }
    if (b0 + 64u * kKeepBatches < n) load_list(b0 + 64u * kKeepBatches);
    bool flag[kKeepBatches];
#pragma unroll
    for (int u = 0; u < kKeepBatches; ++u)
      flag[u] = sid[u] != kNone && ((bits[sid[u] >> 5] >> (sid[u] & 31)) & 1u);
    if (PACKED) {
      // packed entries: the position of the kept ones from their sub-cluster's z-pool offset
#pragma unroll
      for (int u = 0; u < kKeepBatches; ++u)
        if (flag[u]) pos[u] = seg_lds[sid[u]] + (pos[u] & ((1u << pk_shift) - 1u));
    }
    if (filt) {
      if (WORDS) {
        uint64_t wv[kKeepBatches];
#pragma unroll
        for (int u = 0; u < kKeepBatches; ++u) wv[u] = lw.w[flag[u] ? pos[u] : 0u];
#pragma unroll
        for (int u = 0; u < kKeepBatches; ++u) flag[u] = flag[u] && label_word_pass(lw, box, wv[u]);
      } else if (l == 4) {
        float4 lv[kKeepBatches];
#pragma unroll
        for (int u = 0; u < kKeepBatches; ++u)
          lv[u] = flag[u] ? reinterpret_cast<const float4*>(lab)[pos[u]] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int u = 0; u < kKeepBatches; ++u)
          flag[u] = flag[u] && !(lv[u].x < rg[0] || lv[u].x > rg[1] || lv[u].y < rg[2] || lv[u].y > rg[3] ||
                                 lv[u].z < rg[4] || lv[u].z > rg[5] || lv[u].w < rg[6] || lv[u].w > rg[7]);
      } else {
#pragma unroll
        for (int u = 0; u < kKeepBatches; ++u) flag[u] = flag[u] && label_pass(lab, l, pos[u], rg);
      }
    }
#pragma unroll
    for (int u = 0; u < kKeepBatches; ++u) {
      const uint64_t mask = __ballot(flag[u]);
      if (flag[u]) out[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = pos[u];
      base += (uint32_t)__popcll(mask);
    }
  }
  if (lane == 0) nL[qi] = base;
}

// A11: ADC distances (product-quantization.cu:37-77) written as f32 per column; the selections read them as
// unique keys (distance bits, column).
// Mode 1: the query's Morton-range candidates, column = candidate column.  Mode 2 (EXPS0 <= 1): the
// label-filtered refine list, column = list index.  Work = chunks of PER * 256 items of a query
// (coff[q] = chunks before query q); the grid is persistent and each workgroup takes a contiguous run of
// chunks, so a query's LUT (product-quantization.cu:4-35, an fma chain over pq_len) is built once per
// workgroup when it fits one LDS phase of kLutPhaseBytes, else rebuilt phase by phase per chunk.  Each
// thread owns PER items (lane-consecutive columns, so code rows are read coalesced) and keeps their
// partial sums in registers; the sum over m stays sequential for any pq_dim.  The query's range table
// sits in LDS for the column -> z-pool position lookup.
constexpr uint32_t kLutPhaseBytes = 48 * 1024;
constexpr uint32_t kAdcLdsRanges = 2040;  // LUT phase + range table stay within 64 KiB of LDS

// LUT entry types: f32 (reference-exact) or f16 (PRECISION 1); sums always accumulate in f32.
__device__ __forceinline__ float lutf(float v) { return v; }
__device__ __forceinline__ float lutf(__half v) { return __half2float(v); }
__device__ __forceinline__ void lut_store(float* p, float v) { *p = v; }
__device__ __forceinline__ void lut_store(__half* p, float v) { *p = __float2half(v); }

template <typename LT>
__device__ __forceinline__ float adc_word(float sum, uint32_t wv, const LT* lr, uint64_t K) {
  sum += lutf(lr[wv & 0xFF]);
  sum += lutf(lr[K + ((wv >> 8) & 0xFF)]);
  sum += lutf(lr[2 * K + ((wv >> 16) & 0xFF)]);
  sum += lutf(lr[3 * K + (wv >> 24)]);
  return sum;
}

// Range of a chunk's distance bits: wave reduction, then one atomic min / max per wave.
__device__ __forceinline__ void sel_range_add(uint32_t* mm, uint32_t mn, uint32_t mx) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t a = (uint32_t)__shfl_xor((int)mn, o, 64), b = (uint32_t)__shfl_xor((int)mx, o, 64);
    mn = a < mn ? a : mn;
    mx = b > mx ? b : mx;
  }
  if ((threadIdx.x & 63) == 0 && mn <= mx) {
    atomicMin(&mm[0], mn);
    atomicMax(&mm[1], mx);
  }
}

// STREAM: 0 = per-thread column state over LUT phases; 32 / 8 = streaming, code rows read in steps of 32 codes
// (16-byte loads, pq_dim % 32 == 0) or 8 codes (8-byte loads, pq_dim % 8 == 0).
template <typename LT, int PER, int KC, int STREAM>
__global__ __launch_bounds__(256) void adc_scan_kernel(
    uint64_t qn, uint64_t q0, const float* __restrict__ queries, uint64_t d, uint64_t M, uint64_t K_rt,
    uint64_t pq_len, const float* __restrict__ pq_centers, const uint8_t* __restrict__ codes,
    uint32_t sub_lists, const uint32_t* __restrict__ mode, const uint32_t* __restrict__ n_chosen_s,
    const uint32_t* __restrict__ rstart, const uint32_t* __restrict__ rpre, const uint64_t* __restrict__ cand,
    const uint32_t* __restrict__ L, const uint64_t* __restrict__ loff, const uint32_t* __restrict__ nL,
    const uint64_t* __restrict__ coff, const uint64_t* __restrict__ aoff, uint32_t* __restrict__ adist,
    uint32_t lut_bytes, uint32_t* __restrict__ smm, const uint32_t* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  extern __shared__ __align__(16) unsigned char smem[];
  const uint64_t K = KC ? (uint64_t)KC : K_rt;  // codebook size; a compile-time 256 folds the LUT row offsets
  LT* lut = reinterpret_cast<LT*>(smem);  // [mp][K]
  // the range table is staged in LDS when it fits beside the LUT (sub_lists <= kAdcLdsRanges)
  const bool table_lds = sub_lists <= kAdcLdsRanges;
  uint32_t* rs_l = reinterpret_cast<uint32_t*>(smem + lut_bytes);
  uint32_t* rp_l = rs_l + sub_lists;
  constexpr uint32_t kChunk = PER * kWG;
  const uint64_t total = coff[qn];
  const uint64_t per_wg = (total + gridDim.x - 1) / gridDim.x;
  uint64_t blk = (uint64_t)blockIdx.x * per_wg;
  const uint64_t blk_end = blk + per_wg < total ? blk + per_wg : total;
  if (blk >= blk_end) return;
  uint64_t mp = kLutPhaseBytes / (K * sizeof(LT));
  if (mp > 4) mp &= ~3ull;
  const bool one_phase = mp >= M;
  const bool words = (M % 4) == 0 && (mp % 4) == 0;
  const bool quads = (M % 16) == 0 && (mp % 16) == 0;
  uint64_t qi = 0;
  {
    uint64_t lo = 0, hi = qn;  // last q with coff[q] <= blk
    while (hi - lo > 1) {
      const uint64_t mid = (lo + hi) >> 1;
      if (coff[mid] <= blk) lo = mid; else hi = mid;
    }
    qi = lo;
  }
  uint64_t cur = ~0ull;
  for (; blk < blk_end; ++blk) {
    while (coff[qi + 1] <= blk) ++qi;
    const bool from_list = mode[qi] == 2;
    const uint32_t n_items = (uint32_t)(from_list ? nL[qi] : cand[qi]);
    const uint32_t c0 = (uint32_t)(blk - coff[qi]) * kChunk;
    const uint32_t c1 = c0 + kChunk < n_items ? c0 + kChunk : n_items;
    const float* q = queries + (q0 + qi) * d;
    const uint32_t nr = n_chosen_s[qi];
    const uint32_t* list = L + loff[qi];
    const bool fresh = qi != cur;
    const uint32_t* rs = table_lds ? rs_l : rstart + qi * sub_lists;
    const uint32_t* rp = table_lds ? rp_l : rpre + qi * (sub_lists + 1);
    __syncthreads();  // the previous chunk is done with the LUT and the range table
    if (fresh && !from_list && table_lds) {
      for (uint32_t j = threadIdx.x; j <= nr; j += kWG) {
        rp_l[j] = rpre[qi * (sub_lists + 1) + j];
        if (j < nr) rs_l[j] = rstart[qi * sub_lists + j];
      }
    }
    if (fresh && one_phase) {
      for (uint64_t e = threadIdx.x; e < M * K; e += kWG) {
        const uint64_t m = e / K, c = e % K;
        float ans = 0.0f;
        for (uint64_t t = 0; t < pq_len; ++t) {
          if (m * pq_len + t >= d) break;
          const float diff = pq_centers[m * pq_len * K + c * pq_len + t] - q[m * pq_len + t];
          ans = fmaf(diff, diff, ans);
        }
        lut_store(&lut[e], ans);
      }
    }
    cur = qi;
    __syncthreads();
    if constexpr (STREAM != 0) {
      // streaming (the host picks it when the LUT fits one phase): groups of 4 of this thread's columns, each
      // finished (all M codes) before the next, so no per-column state stays live and occupancy stays high
      uint32_t j = 0;
      if (!from_list) {
        const uint32_t col = c0 + threadIdx.x;
        uint32_t hi = nr;
        while (hi - j > 1) {
          const uint32_t mid = (j + hi) >> 1;
          if (rp[mid] <= col) j = mid; else hi = mid;
        }
      }
      uint32_t* out = adist + aoff[qi];
      uint32_t tmn = 0xFFFFFFFFu, tmx = 0u;
#pragma unroll 1
      for (int i0 = 0; i0 < PER; i0 += 4) {
        uint32_t p4[4];
        float sum[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t col = c0 + threadIdx.x + (uint32_t)(i0 + u) * kWG;
          sum[u] = 0.0f;
          p4[u] = 0;
          if (col < c1) {
            if (from_list) {
              p4[u] = list[col];
            } else {
              while (j + 1 < nr && rp[j + 1] <= col) ++j;
              p4[u] = rs[j] + (col - rp[j]);
            }
          }
        }
        if (c0 + threadIdx.x + (uint32_t)i0 * kWG >= c1) break;  // the rest of this thread's columns too
        if constexpr (STREAM == 8) {
          // a code row of 8, 16 or 24 codes: one 16-byte word and/or one 8-byte tail from the 64-row blocked
          // codes, all four rows' pieces in flight together; the sum keeps the order over m
          uint4 a[4];
          uint2 b[4];
          const bool has16 = M >= 16, has8 = (M & 15) != 0;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint32_t col = c0 + threadIdx.x + (uint32_t)(i0 + u) * kWG;
            const uint64_t p = p4[u];
            const uint8_t* blk = codes + (p >> 6) * 64 * M;
            const bool ok = col < c1;
            a[u] = ok && has16 ? reinterpret_cast<const uint4*>(blk)[p & 63] : make_uint4(0, 0, 0, 0);
            b[u] = ok && has8 ? reinterpret_cast<const uint2*>(blk + (M / 16) * 1024)[p & 63] : make_uint2(0, 0);
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const LT* lr = lut;
            if (has16) {
              sum[u] = adc_word(sum[u], a[u].x, lr, K);
              sum[u] = adc_word(sum[u], a[u].y, lr + 4 * K, K);
              sum[u] = adc_word(sum[u], a[u].z, lr + 8 * K, K);
              sum[u] = adc_word(sum[u], a[u].w, lr + 12 * K, K);
              lr += 16 * K;
            }
            if (has8) {
              sum[u] = adc_word(sum[u], b[u].x, lr, K);
              sum[u] = adc_word(sum[u], b[u].y, lr + 4 * K, K);
            }
          }
        } else {
        // codes in 64-row blocks of 16-byte words: word w of row p at block (p / 64), lane slot p % 64
#pragma unroll 1
        for (uint64_t mb = 0; mb < M; mb += 32) {
          uint4 v[4][2];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint32_t col = c0 + threadIdx.x + (uint32_t)(i0 + u) * kWG;
            const uint64_t p = p4[u];
            const uint4* cd = reinterpret_cast<const uint4*>(codes + (p >> 6) * 64 * M) + (mb / 16) * 64 + (p & 63);
            const bool ok = col < c1;
            v[u][0] = ok ? cd[0] : make_uint4(0, 0, 0, 0);
            v[u][1] = ok ? cd[64] : make_uint4(0, 0, 0, 0);
          }
          const LT* lr0 = lut + mb * K;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              const LT* lr = lr0 + (uint64_t)h * 16 * K;
              sum[u] = adc_word(sum[u], v[u][h].x, lr, K);
              sum[u] = adc_word(sum[u], v[u][h].y, lr + 4 * K, K);
              sum[u] = adc_word(sum[u], v[u][h].z, lr + 8 * K, K);
              sum[u] = adc_word(sum[u], v[u][h].w, lr + 12 * K, K);
            }
          }
        }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t col = c0 + threadIdx.x + (uint32_t)(i0 + u) * kWG;
          if (col < c1) {
            const uint32_t b = __float_as_uint(sum[u]);
            out[col] = b;
            tmn = b < tmn ? b : tmn;
            tmx = b > tmx ? b : tmx;
          }
        }
      }
      if (!from_list) sel_range_add(smm + 2 * qi, tmn, tmx);
      continue;
    } else {
    uint32_t pos[PER];
    float part[PER];
    if (from_list) {
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const uint32_t col = c0 + threadIdx.x + (uint32_t)i * kWG;
        part[i] = 0.0f;
        pos[i] = col < c1 ? list[col] : 0u;
      }
    } else {
      // this thread's columns ascend by kWG: one search for the first, then step through the ranges
      uint32_t j = 0;
      {
        const uint32_t col = c0 + threadIdx.x;
        uint32_t hi = nr;
        while (hi - j > 1) {
          const uint32_t mid = (j + hi) >> 1;
          if (rp[mid] <= col) j = mid; else hi = mid;
        }
      }
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const uint32_t col = c0 + threadIdx.x + (uint32_t)i * kWG;
        part[i] = 0.0f;
        pos[i] = 0;
        if (col < c1) {
          while (j + 1 < nr && rp[j + 1] <= col) ++j;
          pos[i] = rs[j] + (col - rp[j]);
        }
      }
    }
    for (uint64_t m0 = 0; m0 < M; m0 += mp) {
      const uint64_t m1 = m0 + mp < M ? m0 + mp : M;
      if (!one_phase) {
        if (m0 > 0) __syncthreads();
        for (uint64_t e = threadIdx.x; e < (m1 - m0) * K; e += kWG) {
          const uint64_t m = m0 + e / K, c = e % K;
          float ans = 0.0f;
          for (uint64_t t = 0; t < pq_len; ++t) {
            if (m * pq_len + t >= d) break;
            const float diff = pq_centers[m * pq_len * K + c * pq_len + t] - q[m * pq_len + t];
            ans = fmaf(diff, diff, ans);
          }
          lut_store(&lut[e], ans);
        }
        __syncthreads();
      }
      if (quads && ((m1 - m0) & 31) == 0) {
        // 32 codes (two 16-byte pieces) per item and step; four items' pieces are in flight before their
        // sums start; each item still adds its m in ascending order
#pragma unroll
        for (int i0 = 0; i0 < PER; i0 += 4) {
          for (uint64_t mb = m0; mb < m1; mb += 32) {
            uint4 v[4][2];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const uint32_t col = c0 + threadIdx.x + (uint32_t)(i0 + u) * kWG;
              const uint4* cd = reinterpret_cast<const uint4*>(codes + (uint64_t)pos[i0 + u] * M + mb);
              const bool ok = col < c1;
              v[u][0] = ok ? cd[0] : make_uint4(0, 0, 0, 0);
              v[u][1] = ok ? cd[1] : make_uint4(0, 0, 0, 0);
            }
            const LT* lr0 = lut + (mb - m0) * K;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              float sum = part[i0 + u];
#pragma unroll
              for (int h = 0; h < 2; ++h) {
                const LT* lr = lr0 + (uint64_t)h * 16 * K;
                sum = adc_word(sum, v[u][h].x, lr, K);
                sum = adc_word(sum, v[u][h].y, lr + 4 * K, K);
                sum = adc_word(sum, v[u][h].z, lr + 8 * K, K);
                sum = adc_word(sum, v[u][h].w, lr + 12 * K, K);
              }
              part[i0 + u] = sum;
            }
          }
        }
        continue;
      }
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const uint32_t col = c0 + threadIdx.x + (uint32_t)i * kWG;
        if (col >= c1) continue;
        const uint8_t* cd = codes + (uint64_t)pos[i] * M;
        float sum = part[i];
        if (quads) {
          for (uint64_t m = m0; m < m1; m += 16) {
            const uint4 v = *reinterpret_cast<const uint4*>(cd + m);
            const LT* lr = lut + (m - m0) * K;
            sum = adc_word(sum, v.x, lr, K);
            sum = adc_word(sum, v.y, lr + 4 * K, K);
            sum = adc_word(sum, v.z, lr + 8 * K, K);
            sum = adc_word(sum, v.w, lr + 12 * K, K);
          }
        } else if (words) {
          for (uint64_t m = m0; m < m1; m += 4)
            sum = adc_word(sum, *reinterpret_cast<const uint32_t*>(cd + m), lut + (m - m0) * K, K);
        } else {
          for (uint64_t m = m0; m < m1; ++m) sum += lutf(lut[(m - m0) * K + cd[m]]);
        }
        part[i] = sum;
      }
    }
    uint32_t* out = adist + aoff[qi];
    uint32_t tmn = 0xFFFFFFFFu, tmx = 0u;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const uint32_t col = c0 + threadIdx.x + (uint32_t)i * kWG;
      if (col < c1) {
        const uint32_t b = __float_as_uint(part[i]);
        out[col] = b;
        tmn = b < tmn ? b : tmn;
        tmx = b > tmx ? b : tmx;
      }
    }
    if (!from_list) sel_range_add(smm + 2 * qi, tmn, tmx);
    }
  }
}

// The tile of this workgroup over the concatenated ADC distances, cut at query boundaries; f(qi, a, e) runs
// for the candidates [a, e) of each query with a selection slot (uniform over the workgroup).
template <class F>
__device__ inline void for_each_sel_segment(uint64_t qn, const uint64_t* __restrict__ aoff,
                                            const uint32_t* __restrict__ sslot, F f) {
  const uint64_t total = aoff[qn];
  const uint64_t t0 = (uint64_t)blockIdx.x * kSelTile;
  if (t0 >= total) return;
  const uint64_t t1 = t0 + kSelTile < total ? t0 + kSelTile : total;
  uint64_t lo = 0, hi = qn;  // last q with aoff[q] <= t0
  while (hi - lo > 1) {
    const uint64_t mid = (lo + hi) >> 1;
    if (aoff[mid] <= t0) lo = mid; else hi = mid;
  }
  uint64_t qi = lo;
  for (uint64_t x = t0; x < t1;) {
    while (aoff[qi + 1] <= x) ++qi;
    const uint64_t e = aoff[qi + 1] < t1 ? aoff[qi + 1] : t1;
    if (sslot[qi] != kNone) f(qi, (uint32_t)(x - aoff[qi]), (uint32_t)(e - aoff[qi]));
    x = e;
  }
}

// Histogram pass of the many-workgroup selection: LDS counts of a tile, added into the slot's histogram.
__global__ __launch_bounds__(256) void sel_hist_kernel(uint64_t qn, const uint64_t* __restrict__ aoff,
                                                       const uint32_t* __restrict__ sslot,
                                                       const uint32_t* __restrict__ smm,
                                                       const uint32_t* __restrict__ adist, uint32_t* __restrict__ ghist,
                                                       const uint32_t* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  __shared__ __align__(16) uint32_t hist[kSelBins];
  constexpr uint32_t kStep = kWG * dev::kSelUnroll;
  for_each_sel_segment(qn, aoff, sslot, [&](uint64_t qi, uint32_t a, uint32_t e) {
    const uint32_t mn = smm[2 * qi], sft = dev::dist_bin_shift(mn, smm[2 * qi + 1]);
    for (uint32_t i = threadIdx.x; i < kSelBins; i += kWG) hist[i] = 0;
    __syncthreads();
    const uint32_t* d = adist + aoff[qi];
    for (uint32_t i0 = a + threadIdx.x; i0 < e; i0 += kStep) {
      uint32_t v[dev::kSelUnroll];
#pragma unroll
      for (int u = 0; u < dev::kSelUnroll; ++u) {
        const uint32_t i = i0 + (uint32_t)u * kWG;
        v[u] = i < e ? d[i] : 0u;
      }
#pragma unroll
      for (int u = 0; u < dev::kSelUnroll; ++u)
        if (i0 + (uint32_t)u * kWG < e) atomicAdd(&hist[(v[u] - mn) >> sft], 1u);
    }
    __syncthreads();
    uint32_t* g = ghist + (uint64_t)sslot[qi] * kSelBins;
    for (uint32_t i = threadIdx.x; i < kSelBins; i += kWG) {
      const uint32_t c = hist[i];
      if (c) atomicAdd(&g[i], c);
    }
    __syncthreads();
  });
}

// Items per thread and step of the split (8 would cost occupancy: the label loads hold many registers).
constexpr int kSplitUnroll = 4;
static_assert(kSelTile <= kWG * 64, "a tile's items per thread fit the 64-bit masks of the split");

// The split of a query's candidates [a, e) (at most 64 per thread; distance bits d[], bins
// (bits - mn) >> sft) around the k'-th key's bin bstar, with the A12 post-filter folded in.  Candidates in
// lower bins are in the top k' for sure: those passing the label filter (label_pass) go straight to the
// refine list out[] as z-pool positions (a thread's columns ascend, so it walks the query's range table
// rs/rp forward, and a wave's label reads are contiguous within a range).  Candidates in bin bstar go to
// scr[] for the final ranking.  A first pass marks both kinds in per-thread bit masks, one workgroup scan and
// one reservation on the global counters cnt[0] (survivors) / cnt[1] (bin keys) place them, a second pass
// writes them.  rg: the query's 2 l label bounds.  Called by the whole workgroup.
template <bool WORDS>
__device__ inline void sel_split_range(uint32_t a, uint32_t e, const uint32_t* __restrict__ d, uint32_t mn,
                                       uint32_t sft, uint32_t bstar, uint32_t nr, const uint32_t* rs,
                                       const uint32_t* rp, int l, const float* __restrict__ lab,
                                       const LabelWords& lw, const float* rg, uint32_t* __restrict__ out,
                                       uint64_t* __restrict__ scr, uint32_t* cnt) {
  __shared__ uint32_t tmp[kWG / 64 + 1];
  __shared__ uint32_t res[2];
  constexpr uint32_t kStep = kWG * kSplitUnroll;
  uint32_t j0 = 0;  // range of this thread's first column
  {
    const uint32_t col = a + threadIdx.x;
    uint32_t hi = nr;
    while (hi - j0 > 1) {
      const uint32_t mid = (j0 + hi) >> 1;
      if (rp[mid] <= col) j0 = mid; else hi = mid;
    }
  }
  uint32_t j = j0;
  const LabelBox box = WORDS && l > 1 ? label_box(lw, l, rg) : LabelBox{};
  uint64_t flo = 0, fin = 0;  // bit step * kSplitUnroll + u: listed survivor / key in the bin
  uint32_t step = 0;
  // the next step's distances load while this step's label loads are in flight (indices clamped into [a, e),
  // so every load is issued unconditionally)
  uint32_t vn[kSplitUnroll];
  auto load_d = [&](uint32_t base) {
#pragma unroll
    for (int u = 0; u < kSplitUnroll; ++u) {
      const uint32_t i = base + (uint32_t)u * kWG + threadIdx.x;
      vn[u] = d[i < e ? i : e - 1];
    }
  };
  if (a < e) load_d(a);
  for (uint32_t base = a; base < e; base += kStep, ++step) {
    uint32_t v[kSplitUnroll], pos[kSplitUnroll];
#pragma unroll
    for (int u = 0; u < kSplitUnroll; ++u) v[u] = vn[u];
    if (base + kStep < e) load_d(base + kStep);
    uint32_t lo = 0;  // items of this step below the bin
#pragma unroll
    for (int u = 0; u < kSplitUnroll; ++u) {
      const uint32_t i = base + (uint32_t)u * kWG + threadIdx.x;
      pos[u] = 0;
      if (i < e) {
        const uint32_t b = (v[u] - mn) >> sft;
        if (b < bstar) {
          while (j + 1 < nr && rp[j + 1] <= i) ++j;
          pos[u] = rs[j] + (i - rp[j]);
          lo |= 1u << u;
        } else if (b == bstar) {
          fin |= 1ull << (step * kSplitUnroll + u);
        }
      }
    }
    // the label filter, every item's loads in flight together
    if (WORDS) {
      if (l > 1) {
      uint64_t wv[kSplitUnroll];
#pragma unroll
      for (int u = 0; u < kSplitUnroll; ++u) wv[u] = lw.w[pos[u]];  // pos 0 where not below the bin
#pragma unroll
      for (int u = 0; u < kSplitUnroll; ++u)
        if (!label_word_pass(lw, box, wv[u])) lo &= ~(1u << u);
      }
    } else if (l == 4) {
      float4 lv[kSplitUnroll];
#pragma unroll
      for (int u = 0; u < kSplitUnroll; ++u)
        lv[u] = (lo >> u) & 1u ? reinterpret_cast<const float4*>(lab)[pos[u]] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (int u = 0; u < kSplitUnroll; ++u) {
        const bool fail = lv[u].x < rg[0] || lv[u].x > rg[1] || lv[u].y < rg[2] || lv[u].y > rg[3] ||
                          lv[u].z < rg[4] || lv[u].z > rg[5] || lv[u].w < rg[6] || lv[u].w > rg[7];
        if (fail) lo &= ~(1u << u);
      }
    } else if (l > 1) {
      for (int dm = 0; dm < l; ++dm) {
        float lv[kSplitUnroll];
#pragma unroll
        for (int u = 0; u < kSplitUnroll; ++u) lv[u] = (lo >> u) & 1u ? lab[(uint64_t)pos[u] * l + dm] : 0.f;
        const float lo_b = rg[2 * dm], hi_b = rg[2 * dm + 1];
#pragma unroll
        for (int u = 0; u < kSplitUnroll; ++u)
          if (lv[u] < lo_b || lv[u] > hi_b) lo &= ~(1u << u);
      }
    }
    flo |= (uint64_t)lo << (step * kSplitUnroll);
  }
  // places: survivors in the low 16 bits, bin keys in the high 16 bits (a segment holds at most 16384)
  const uint32_t c = (uint32_t)__popcll(flo) | ((uint32_t)__popcll(fin) << 16);
  uint32_t total;
  const uint32_t excl = dev::wg_exclusive_scan(c, tmp, total);
  if (threadIdx.x == 0) {
    res[0] = (total & 0xFFFFu) ? atomicAdd(&cnt[0], total & 0xFFFFu) : 0u;
    res[1] = (total >> 16) ? atomicAdd(&cnt[1], total >> 16) : 0u;
  }
  __syncthreads();
  uint32_t olo = res[0] + (excl & 0xFFFFu), oin = res[1] + (excl >> 16);
  j = j0;
  while (flo) {
    const int t = __ffsll((long long)flo) - 1;
    flo &= flo - 1;
    const uint32_t i = a + (uint32_t)(t / kSplitUnroll) * kStep + (uint32_t)(t % kSplitUnroll) * kWG + threadIdx.x;
    while (j + 1 < nr && rp[j + 1] <= i) ++j;
    out[olo++] = rs[j] + (i - rp[j]);
  }
  while (fin) {
    const int t = __ffsll((long long)fin) - 1;
    fin &= fin - 1;
    const uint32_t i = a + (uint32_t)(t / kSplitUnroll) * kStep + (uint32_t)(t % kSplitUnroll) * kWG + threadIdx.x;
    scr[oin++] = ((uint64_t)d[i] << 32) | i;
  }
}

// Stages a query's range table (nr <= kSelLdsRanges) and label bounds in LDS.
constexpr uint32_t kSelLdsRanges = 1024;

__device__ inline void sel_stage_query(uint64_t qi, uint32_t nr, uint32_t sub_lists, int l,
                                       const uint32_t* __restrict__ rstart, const uint32_t* __restrict__ rpre,
                                       const float* __restrict__ ranges, uint32_t* rs_l, uint32_t* rp_l, float* rg) {
  if (nr <= kSelLdsRanges) {
    for (uint32_t j = threadIdx.x; j <= nr; j += kWG) {
      rp_l[j] = rpre[qi * (sub_lists + 1) + j];
      if (j < nr) rs_l[j] = rstart[qi * sub_lists + j];
    }
  }
  if (threadIdx.x < 2 * (uint32_t)l) rg[threadIdx.x] = ranges[qi * 2 * l + threadIdx.x];
}

// Split pass of the many-workgroup selection over tiles.  Per query, scnt[4 qi + 0] counts the listed
// survivors, [1] the keys in the k'-th key's bin and [2] holds the keys below that bin.
template <bool WORDS>
__global__ __launch_bounds__(256) void sel_split_kernel(
    uint64_t qn, int l, uint32_t sub_lists, const uint64_t* __restrict__ aoff, const uint32_t* __restrict__ sslot,
    const uint32_t* __restrict__ smm, const uint64_t* __restrict__ kprime, const uint32_t* __restrict__ adist,
    const uint32_t* __restrict__ ghist, const uint32_t* __restrict__ n_chosen_s, const uint32_t* __restrict__ rstart,
    const uint32_t* __restrict__ rpre, const float* __restrict__ lab, LabelWords lw,
    const float* __restrict__ ranges, const uint64_t* __restrict__ loff, uint32_t* __restrict__ L,
    uint64_t* __restrict__ scratch, uint32_t* __restrict__ scnt, const uint32_t* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  __shared__ __align__(16) uint32_t hist[kSelBins];
  __shared__ uint32_t rs_l[kSelLdsRanges], rp_l[kSelLdsRanges + 1];
  __shared__ float rg[2 * kMaxDims];
  __shared__ uint64_t sh[8];
  for_each_sel_segment(qn, aoff, sslot, [&](uint64_t qi, uint32_t a, uint32_t e) {
    const uint32_t mn = smm[2 * qi], sft = dev::dist_bin_shift(mn, smm[2 * qi + 1]);
    const uint32_t nr = n_chosen_s[qi];
    sel_stage_query(qi, nr, sub_lists, l, rstart, rpre, ranges, rs_l, rp_l, rg);
    const uint4* g = reinterpret_cast<const uint4*>(ghist + (uint64_t)sslot[qi] * kSelBins);
    for (uint32_t i = threadIdx.x; i < kSelBins / 4; i += kWG) reinterpret_cast<uint4*>(hist)[i] = g[i];
    __syncthreads();
    dev::wg_hist_find(hist, (uint32_t)kprime[qi], sh);
    if (threadIdx.x == 0) scnt[4 * qi + 2] = (uint32_t)sh[1];  // the same value from every tile of the query
    const bool table_lds = nr <= kSelLdsRanges;
    sel_split_range<WORDS>(a, e, adist + aoff[qi], mn, sft, (uint32_t)sh[0], nr,
                           table_lds ? rs_l : rstart + qi * sub_lists, table_lds ? rp_l : rpre + qi * (sub_lists + 1),
                           l, lab, lw, rg, L + loff[qi], scratch + aoff[qi], scnt + 4 * qi);
    __syncthreads();
  });
}

// Clears the global histograms of the slots handed out by the plan.
__global__ void sel_zero_kernel(const uint32_t* __restrict__ slot_ctr, uint32_t slot_cap, uint32_t* __restrict__ ghist) {
  const uint32_t used = *slot_ctr < slot_cap ? *slot_ctr : slot_cap;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (uint64_t)used * (kSelBins / 4)) reinterpret_cast<uint4*>(ghist)[i] = make_uint4(0, 0, 0, 0);
}

constexpr int kOneWgUnroll = 2;  // distance loads in flight per thread of the one-workgroup selection

// A11 selection + A12 post-filter for mode-1 queries: the k' smallest (pq distance, column) keys
// (select_k, product-quantization.cu:264), then for l > 1 the exact label filter.  A query with a selection
// slot arrives split by the tiles (survivors below the k'-th key's bin already listed); the others are
// selected here by one workgroup.
__global__ __launch_bounds__(256) void adc_select_kernel(
    uint64_t qn, int l, uint32_t sub_lists, const uint32_t* __restrict__ mode,
    const uint64_t* __restrict__ kprime, const uint64_t* __restrict__ cand,
    const uint32_t* __restrict__ n_chosen_s, const uint32_t* __restrict__ rstart,
    const uint32_t* __restrict__ rpre, const float* __restrict__ lab, LabelWords lw,
    const float* __restrict__ ranges, const uint64_t* __restrict__ aoff, uint64_t* __restrict__ akeys_sel_scratch,
    const uint32_t* __restrict__ adist, uint64_t* __restrict__ akeys2, const uint64_t* __restrict__ loff,
    uint32_t* __restrict__ L, uint32_t* __restrict__ nL, const uint32_t* __restrict__ sslot,
    const uint32_t* __restrict__ smm, const uint32_t* __restrict__ scnt, const uint32_t* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  __shared__ __align__(16) uint32_t hist[kSelBins];
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
  const uint32_t nr = n_chosen_s[qi];
  const uint32_t* rs = rstart + qi * sub_lists;
  const uint32_t* rp = rpre + qi * (sub_lists + 1);
  const uint64_t* sel = akeys_sel_scratch + loff[qi];  // the selected keys still to post-filter
  uint32_t nsel = kp, listed = 0;                        // their number; refine-list entries already written
  if (sslot[qi] != kNone) {
    // the tiles have listed the survivors below the k'-th key's bin: rank that bin's keys
    listed = scnt[4 * qi];
    const uint32_t m = scnt[4 * qi + 1];
    nsel = kp - scnt[4 * qi + 2];
    const uint64_t* scr = akeys2 + aoff[qi];
    if (nsel == m) {
      sel = scr;
    } else {
      dev::wg_radix_select(dev::KeyArray{scr}, m, nsel, akeys_sel_scratch + loff[qi], hist, sh);
    }
  } else {
    dev::wg_dist_select<kOneWgUnroll>(adist + aoff[qi], Cq, kp, smm[2 * qi], smm[2 * qi + 1], akeys_sel_scratch + loff[qi],
                        akeys2 + aoff[qi], hist, sh);
  }
  float rg[2 * kMaxDims];
  for (int i = 0; i < 2 * l; ++i) rg[i] = ranges[qi * 2 * l + i];
  const LabelBox box = lw.w && l > 1 ? label_box(lw, l, rg) : LabelBox{};
  if (threadIdx.x == 0) cnt_sh = 0;
  // the range table in LDS (the histogram's space, free after the selection) for the column lookups
  const bool table_lds = 2 * nr + 1 <= (uint32_t)kSelBins;
  const uint32_t* rs_c = rs;
  const uint32_t* rp_c = rp;
  if (table_lds) {
    __syncthreads();  // the selection is done with the histogram
    for (uint32_t j = threadIdx.x; j <= nr; j += kWG) {
      hist[nr + j] = rp[j];
      if (j < nr) hist[j] = rs[j];
    }
    rs_c = hist;
    rp_c = hist + nr;
  }
  __syncthreads();
  uint32_t* out = L + loff[qi] + listed;
  for (uint32_t e = threadIdx.x; e < nsel; e += kWG) {
    const uint32_t pos = col_to_pos(rs_c, rp_c, nr, dev::key_tie(sel[e]));
    if (l == 1) {
      out[e] = pos;
    } else if (lw.w ? label_word_pass(lw, box, lw.w[pos]) : label_pass(lab, l, pos, rg)) {
      out[atomicAdd(&cnt_sh, 1u)] = pos;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) nL[qi] = listed + (l == 1 ? nsel : cnt_sh);
}

// A13 select: top-k by (exact distance, id) of the refine list with the reference's pad columns
// (refine.cu:78-119), or for the query_prefilter path the top-k by (pq distance, column).
__global__ __launch_bounds__(256) void final_pick_kernel(
    uint64_t q0, uint32_t scap, const uint32_t* __restrict__ zid, int l, uint64_t topk, const uint32_t* __restrict__ mode,
    const uint64_t* __restrict__ kprime, const uint64_t* __restrict__ loff, const uint32_t* __restrict__ L,
    const uint32_t* __restrict__ nL, const uint64_t* __restrict__ pad32, const uint64_t* __restrict__ padmax,
    uint64_t* __restrict__ keys, const uint32_t* __restrict__ adist, const uint64_t* __restrict__ aoff,
    float* __restrict__ out_dist, uint64_t* __restrict__ out_ids, const uint32_t* __restrict__ abort_flag,
    int64_t skip_upto) {
  if (abort_flag && *abort_flag) return;
  extern __shared__ __align__(16) unsigned char smem[];
  uint64_t* sorted = reinterpret_cast<uint64_t*>(smem);
  __shared__ __align__(16) uint32_t hist[2048];
  __shared__ uint64_t sh[8];
  const uint64_t qi = blockIdx.x;
  const uint32_t n = nL[qi];
  if (n <= skip_upto) return;  // handled by final_pick_wave_kernel
  const uint32_t* list = L + loff[qi];

  float* od = out_dist + (q0 + qi) * topk;
  uint64_t* oi = out_ids + (q0 + qi) * topk;
  const uint32_t k = (uint32_t)topk;
  const uint32_t m = mode[qi] == 2 ? select_sorted(dev::KeyDist{adist + aoff[qi]}, n, k, sorted, scap, hist, sh)
                                   : select_sorted(dev::KeyArray{keys + loff[qi]}, n, k, sorted, scap, hist, sh);
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
  // pads of the refine list: (1e6, 2^32) x p32 then (1e6, UINT64_MAX) x pmax, placed after the real
  // items of value <= 1e6
  uint64_t p32 = pad32[qi], pmax = padmax[qi];
  if (l > 1) pmax = kprime[qi] > n ? kprime[qi] - n : 0;
  const uint32_t n_le = count_le_pad(sorted, m);
  for (uint32_t j = threadIdx.x; j < k; j += kWG) {
    float dv;
    uint64_t iv;
    if (j < n_le) {
      dv = dev::key_value(sorted[j]);
      iv = dev::key_tie(sorted[j]);
    } else if ((uint64_t)(j - n_le) < p32) {
      dv = kExactPad;
      iv = kIdPad;
    } else if ((uint64_t)(j - n_le) < p32 + pmax) {
      dv = kExactPad;
      iv = kU64Max;
    } else {
      const uint64_t r = (uint64_t)j - p32 - pmax;  // real items beyond 1e6
      if (r < m) {
        dv = dev::key_value(sorted[r]);
        iv = dev::key_tie(sorted[r]);
      } else {
        dv = kExactPad;
        iv = kU64Max;
      }
    }
    od[j] = dv;
    oi[j] = iv;
  }
}

// final_pick_kernel for refine lists of at most kWavePickMax items, one wave per query: the list's keys are
// sorted in registers (the same keys and order as the workgroup sort) and every lane writes its own sorted
// keys to their output slots (shifted past the pad columns), then the pads.  Longer lists are left to
// final_pick_kernel.
constexpr uint32_t kWavePickMax = 512;

template <int R, class Src>
__device__ void wave_pick(Src src, uint32_t n, uint32_t k, int lane, int l, bool adc_mode, uint64_t p32,
                          uint64_t pmax, const uint32_t* __restrict__ zid, const uint32_t* __restrict__ list,
                          float* __restrict__ od, uint64_t* __restrict__ oi) {
  uint64_t key[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)lane * R + r;
    key[r] = e < n ? src(e) : ~0ull;
  }
  dev::wave_bitonic<R>(key, lane);
  const uint32_t m = n < k ? n : k;
  if (adc_mode) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t e = (uint32_t)lane * R + r;
      if (e < m) {
        od[e] = dev::key_value(key[r]);
        oi[e] = zid[list[dev::key_tie(key[r])]];
      }
    }
    for (uint32_t j = m + (uint32_t)lane; j < k; j += 64) {
      od[j] = FLT_MAX;
      oi[j] = kU64Max;
    }
    return;
  }
  // items of value <= 1e6 first, then the pad columns, then the remaining real items
  uint32_t le = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)lane * R + r;
    le += (e < m && dev::key_value(key[r]) <= kExactPad) ? 1u : 0u;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) le += __shfl_xor(le, o, 64);
  const uint32_t n_le = le;
  const uint64_t npad = p32 + pmax;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)lane * R + r;
    if (e >= m) continue;
    const uint64_t j = e < n_le ? (uint64_t)e : (uint64_t)e + npad;
    if (j < k) {
      od[j] = dev::key_value(key[r]);
      oi[j] = dev::key_tie(key[r]);
    }
  }
  for (uint64_t j = (uint64_t)n_le + (uint64_t)lane; j < k; j += 64) {
    const uint64_t t = j - n_le;
    if (t < p32) {
      od[j] = kExactPad;
      oi[j] = kIdPad;
    } else if (t < npad) {
      od[j] = kExactPad;
      oi[j] = kU64Max;
    } else if (j - npad >= m) {
      od[j] = kExactPad;  // beyond the real items
      oi[j] = kU64Max;
    }
  }
}

__global__ __launch_bounds__(256) void final_pick_wave_kernel(
    uint64_t q0, uint64_t qn, const uint32_t* __restrict__ zid, int l, uint64_t topk,
    const uint32_t* __restrict__ mode, const uint64_t* __restrict__ kprime, const uint64_t* __restrict__ loff,
    const uint32_t* __restrict__ L, const uint32_t* __restrict__ nL, const uint64_t* __restrict__ pad32,
    const uint64_t* __restrict__ padmax, const uint64_t* __restrict__ keys, const uint32_t* __restrict__ adist,
    const uint64_t* __restrict__ aoff, float* __restrict__ out_dist, uint64_t* __restrict__ out_ids,
    const uint32_t* __restrict__ abort_flag, uint32_t* __restrict__ left, uint32_t* __restrict__ abort_on_left,
    const uint64_t* __restrict__ cand, const uint64_t* __restrict__ real, const uint32_t* __restrict__ nnz_c,
    const uint32_t* __restrict__ nnz_s, uint64_t* __restrict__ st) {
  if (abort_flag && *abort_flag) return;
  const int lane = threadIdx.x & 63;
  const uint64_t qi = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (qi >= qn) return;
  const uint32_t n = nL[qi];
  if (lane < 6) {
    // the query's counters (as stats_kernel)
    const uint64_t v[6] = {cand[qi], real[qi], n, (uint64_t)nnz_c[qi] + nnz_s[qi], mode[qi] == 1 ? cand[qi] : 0,
                           kprime[qi]};
    st[(q0 + qi) * 6 + lane] = v[lane];
  }
  if (n > kWavePickMax) {
    // left to the workgroup kernel: counted; a replayed graph that left that kernel out is abandoned
    if (lane == 0) {
      atomicAdd(left, 1u);
      if (abort_on_left) *abort_on_left = 1u;
    }
    return;
  }
  const uint32_t* list = L + loff[qi];
  float* od = out_dist + (q0 + qi) * topk;
  uint64_t* oi = out_ids + (q0 + qi) * topk;
  const uint32_t k = (uint32_t)topk;
  const bool adc_mode = mode[qi] == 2;
  uint64_t p32 = 0, pmax = 0;
  if (!adc_mode) {
    p32 = pad32[qi];
    pmax = padmax[qi];
    if (l > 1) pmax = kprime[qi] > n ? kprime[qi] - n : 0;
  }
#define HDG_WAVE_PICK(R)                                                                                      \
  do {                                                                                                        \
    if (adc_mode)                                                                                             \
      wave_pick<R>(dev::KeyDist{adist + aoff[qi]}, n, k, lane, l, true, p32, pmax, zid, list, od, oi);        \
    else                                                                                                      \
      wave_pick<R>(dev::KeyArray{keys + loff[qi]}, n, k, lane, l, false, p32, pmax, zid, list, od, oi);       \
  } while (0)
  if (n <= 64) HDG_WAVE_PICK(1);
  else if (n <= 128) HDG_WAVE_PICK(2);
  else if (n <= 256) HDG_WAVE_PICK(4);
  else HDG_WAVE_PICK(8);
#undef HDG_WAVE_PICK
}

// Per-query counters of the chunk into the call-wide table st[q][6] = (C_q, in-box count of the chosen
// sub-clusters, refined, scored, ADC-scanned, k').
__global__ void stats_kernel(uint64_t qn, uint64_t q0, const uint64_t* __restrict__ cand,
                             const uint64_t* __restrict__ real, const uint32_t* __restrict__ nL,
                             const uint32_t* __restrict__ nnz_c, const uint32_t* __restrict__ nnz_s,
                             const uint32_t* __restrict__ mode, const uint64_t* __restrict__ kprime,
                             uint64_t* __restrict__ st, const uint32_t* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag) return;
  const uint64_t qi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (qi >= qn) return;
  uint64_t* o = st + (q0 + qi) * 6;
  o[0] = cand[qi];
  o[1] = real[qi];
  o[2] = nL[qi];
  o[3] = (uint64_t)nnz_c[qi] + nnz_s[qi];
  o[4] = mode[qi] == 1 ? cand[qi] : 0;
  o[5] = kprime[qi];
}

}  // namespace

// ================================================================================================
// Host orchestration
struct Index::Work {
  DevBuf ranges, boxes, box_lo, box_w, zlo, zhi, zlo_cnt, err;
  DevBuf dense_list, n_dense, counts;                        // A3 dense hand-over
  DevBuf redo;                                               // queries the wave walk hands to the workgroup walk
  DevBuf ib_pos, ib_sub, n_ib;                               // in-box points listed by the walk
  DevBuf sub_list, sub_cnt, ns, list_c, cl_beg, nnz_c;       // non-zero sub-clusters / clusters
  DevBuf chosen_c, n_chosen_c, list_s, list_scnt, nnz_s, chosen_s, n_chosen_s;
  DevBuf keys, woff, bsum, gstat, rstart, rpre, cand, real;
  DevBuf mode, kprime, lsize, asize, chunks, pad32, padmax, nL;
  DevBuf L, Lkeys, adist, akeys2;
  DevBuf sslot, smm, scnt, ghist;  // many-workgroup selection: slots, distance ranges, split counts, histograms
  DevBuf S, feat, budget_T, budget_S, qb16;  // PRECISION 1: scores, model features, budgets, bf16 queries
  DevBuf nzb, nzp;                            // PRECISION 1: non-zero sub-cluster bit words and their ranks
  DevBuf nrng;                         // Morton split mode: range-table pieces in use per query
  DevBuf st, st_chosen;                // call-wide counters [nq][6] and chosen sub-cluster ids [nq][sub_lists]
  uint64_t st_nq = 0, st_sub_lists = 0;
  // GRAPH 1: per-chunk captured graphs, replayed on an internal stream
  struct Captured {
    const void* key[4] = {nullptr, nullptr, nullptr, nullptr};
    uint64_t q0 = 0, qn = 0, G = 0, qcols = 0, cfg = 0, gen = 0;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<Captured> graphs;
  bool capturing = false;
  DevBuf abort_flag;
  // Capacities and fallback kernels of the captured graphs, from the uncaptured runs under the configuration
  // (digest cap_cfg) they are captured for: runs under another configuration (e.g. the plain-path labelling of
  // a budget fit) do not size them
  uint64_t cap_cfg = 0;
  uint64_t lcap = 0, acap = 0;  // refine-list and ADC capacities of the captured graphs
  uint32_t need_mask = 0;       // fallback kernels some sizing run needed (bits as the capture's skip mask)
  uint64_t last_totL = 0, last_totA = 0;  // largest totals seen by uncaptured runs
  hipStream_t gstream = nullptr;
  hipEvent_t gev[2];
  uint64_t* pinned = nullptr;
  bool noalloc = false;  // workspace of query_noalloc: no events, every buffer carved from the scratch array
  hipEvent_t ev[16];
  bool ev_init = false;
  int ev_used = 0;
  const char* const* ev_names = nullptr;
};

// Workspace of the allocation-free query path: one fixed array per device, held until the process ends.  Host
// code never uses this name as an address: the device address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchDevices] = {};

// Device address of the scratch array on `device`, which must be the current device.
void* scratch_address(int device) {
  if (device < 0 || device >= kScratchDevices) fail(kUnsupported, "device ordinal above 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    HDG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail(kHipError, "hipGetSymbolAddress returned a null scratch address");
    g_scratch_ptr[device] = p;
  }
  return g_scratch_ptr[device];
}
// Workgroups of 256 threads of `kernel` (with `lds` bytes of dynamic LDS) resident on the current device at
// once, capped at `want`: the grid of a persistent kernel whose work is split evenly over its workgroups.
unsigned resident_grid(const void* kernel, size_t lds, unsigned want) {
  int dev = 0, ncu = 0, per_cu = 0;
  HDG_HIP(hipGetDevice(&dev));
  HDG_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  HDG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, lds));
  const uint64_t cap = (uint64_t)(per_cu > 0 ? per_cu : 1) * (uint64_t)(ncu > 0 ? ncu : 1);
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(want, cap));
}
}  // namespace

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
  last_w_ = w_;
  HDG_HIP(hipSetDevice(device_));
  HDG_HIP(hipHostMalloc(reinterpret_cast<void**>(&w_->pinned), 64 * sizeof(uint64_t), hipHostMallocDefault));
  // set-up of the allocation-free path: its host-side workspace, pinned read-back words and the scratch address
  wn_ = new Work();
  wn_->noalloc = true;
  HDG_HIP(hipHostMalloc(reinterpret_cast<void**>(&wn_->pinned), 64 * sizeof(uint64_t), hipHostMallocDefault));
  scratch_ = scratch_address(device_);
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
    drop_graphs();
    if (w_->gstream) {
      (void)hipStreamDestroy(w_->gstream);
      for (auto& e : w_->gev) (void)hipEventDestroy(e);
    }
    if (w_->pinned) (void)hipHostFree(w_->pinned);
    if (w_->ev_init)
      for (auto& e : w_->ev) (void)hipEventDestroy(e);
    delete w_;
  }
  if (wn_) {
    if (wn_->pinned) (void)hipHostFree(wn_->pinned);
    delete wn_;  // its buffers are pieces of the scratch array and own nothing
  }
}

void Index::last_stats(uint64_t nq, uint64_t* const out[6]) const {
  Work& w = *last_w_;
  if (nq != w.st_nq) fail(kInvalidArgument, "nq differs from the last query call");
  if (nq == 0) return;
  HDG_HIP(hipSetDevice(device_));
  std::vector<uint64_t> t(nq * 6);
  HDG_HIP(hipMemcpy(t.data(), w.st.data(), nq * 6 * 8, hipMemcpyDeviceToHost));
  for (int c = 0; c < 6; ++c)
    if (out[c])
      for (uint64_t q = 0; q < nq; ++q) out[c][q] = t[q * 6 + c];
}

void Index::last_chosen_subs(uint64_t nq, uint32_t* out) const {
  Work& w = *last_w_;
  if (nq != w.st_nq) fail(kInvalidArgument, "nq differs from the last query call");
  if (nq == 0) return;
  HDG_HIP(hipSetDevice(device_));
  HDG_HIP(hipMemcpy(out, w.st_chosen.data(), nq * w.st_sub_lists * 4, hipMemcpyDeviceToHost));
}

std::vector<std::pair<std::string, float>> Index::stage_times() const {
  Work& w = *last_w_;
  std::vector<std::pair<std::string, float>> r;
  if (w.ev_used < 2) return r;
  HDG_HIP(hipSetDevice(device_));
  HDG_HIP(hipEventSynchronize(w.ev[w.ev_used - 1]));
  for (int i = 0; i + 1 < w.ev_used; ++i) {
    float ms = 0;
    HDG_HIP(hipEventElapsedTime(&ms, w.ev[i], w.ev[i + 1]));
    r.push_back({w.ev_names[i], ms});
  }
  return r;
}

void Index::debug_coarse_scores(const float* queries, uint64_t nq, float* out_host, hipStream_t s) {
  HDG_HIP(hipSetDevice(device_));
  if (!st_.built) fail(kBadState, "query before build");
  DevBuf S, qb;
  S.reserve(nq * st_.C * 4);
  qb.reserve(nq * st_.Dp * 2);
  launch_to_bf16(queries, 0, nq, st_.d, st_.Dp, qb.as<uint16_t>(), s);
  launch_coarse_gemm(qb.as<uint16_t>(), nq, st_.Dp, st_.centers_bf16.as<uint16_t>(), st_.center_norm.as<float>(),
                     (uint32_t)st_.C, S.as<float>(), s);
  HDG_HIP(hipGetLastError());
  HDG_HIP(hipMemcpyAsync(out_host, S.data(), nq * st_.C * 4, hipMemcpyDeviceToHost, s));
  HDG_HIP(hipStreamSynchronize(s));
}

uint64_t Index::check_query(const float* queries, uint64_t nq, uint64_t d, const float* qlabels, uint64_t qcols,
                            float* out_dist, uint64_t* out_ids) const {
  if (!st_.built) fail(kBadState, "query before build");
  if (d != st_.d) fail(kInvalidArgument, "query dimension differs from the index");
  if (nq == 0) return 0;
  if (!queries || !qlabels || !out_dist || !out_ids) fail(kInvalidArgument, "null query/output pointer");
  if (cfg_.is_split && (cfg_.n_split == 0 || cfg_.n_split * cfg_.sub_lists > 8000))
    fail(kUnsupported, "IS_SPLIT = 1 needs N_SPLIT >= 1 and SUB_LISTS * N_SPLIT <= 8000");
  const int l = st_.l;
  if (qcols < (uint64_t)l) fail(kInvalidArgument, "query label columns < l");
  for (int i = 0; i < l; ++i)
    if (fc_.filter_type[i] == 3 && qcols < 2 * (uint64_t)l)
      fail(kInvalidArgument, "explicit-range filter needs 2*l query label columns");
  return cfg_.parity_batch ? cfg_.parity_batch : (uint64_t)find_max_factor(125, nq);
}

void Index::query_noalloc(const float* queries, uint64_t nq, uint64_t d, const float* qlabels, uint64_t qcols,
                          float* out_dist, uint64_t* out_ids, hipStream_t s) {
  HDG_HIP(hipSetDevice(device_));
  const uint64_t G = check_query(queries, nq, d, qlabels, qcols, out_dist, out_ids);
  if (nq == 0) return;
  if (!scratch_) fail(kBadState, "no scratch array on this device");
  if (budget_capturing()) fail(kBadState, "allocation-free query during a budget fit");
  // the call-wide counters and chosen sub-cluster ids stay in the array for the whole call
  const uint64_t call_bytes = ((nq * 6 * 8 + 255) & ~255ull) + ((nq * cfg_.sub_lists * 4 + 255) & ~255ull);
  if (call_bytes > kNoAllocScratchBytes / 4)
    fail(kUnsupported, "allocation-free query: " + std::to_string(nq) + " queries need " +
                           std::to_string(call_bytes) + " bytes of call-wide counters, above a quarter of the " +
                           std::to_string(kNoAllocScratchBytes) + "-byte scratch array; split the call");
  Arena arena{reinterpret_cast<char*>(scratch_), (size_t)kNoAllocScratchBytes, 0};
  Work* const normal = w_;
  struct Restore {
    Index* ix;
    Work* w;
    ~Restore() {
      active_arena() = nullptr;
      ix->w_ = w;
    }
  } restore{this, normal};
  // forget every piece of the previous call (their memory is the same array)
  {
    Work fresh;
    fresh.pinned = wn_->pinned;
    fresh.noalloc = true;
    *wn_ = std::move(fresh);
  }
  w_ = wn_;
  last_w_ = wn_;
  active_arena() = &arena;
  Work& w = *wn_;
  w.st.reserve(nq * 6 * 8);
  w.st_chosen.reserve(nq * cfg_.sub_lists * 4);
  w.st_nq = nq;
  w.st_sub_lists = cfg_.sub_lists;
  w.ev_used = 0;
  const size_t base = arena.used;
  uint64_t chunk = std::max<uint64_t>(G, (cfg_.query_chunk / G) * G);
  for (uint64_t q0 = 0; q0 < nq;) {
    const uint64_t qn = std::min(chunk, nq - q0);
    // every chunk carves its workspace again from the same place; the kernels of the previous chunk are ahead
    // of this one's on the stream, so reusing the memory is ordered
    {
      Work fresh;
      fresh.pinned = wn_->pinned;
      fresh.noalloc = true;
      fresh.st = std::move(wn_->st);
      fresh.st_chosen = std::move(wn_->st_chosen);
      fresh.st_nq = nq;
      fresh.st_sub_lists = cfg_.sub_lists;
      *wn_ = std::move(fresh);
    }
    arena.used = base;
    try {
      query_chunk(queries, q0, qn, qlabels, qcols, G, out_dist, out_ids, s);
    } catch (const ArenaOverflow& e) {
      // halve the chunk (whole reference batches) and run it again; the overwritten outputs of the failed try
      // are written again
      if (qn > G) {
        chunk = std::max<uint64_t>(G, (qn / 2 / G) * G);
        continue;
      }
      fail(kOutOfMemory, "allocation-free query: one reference batch of " + std::to_string(G) +
                             " queries needs " + std::to_string(e.need) + " bytes, above the " +
                             std::to_string(kNoAllocScratchBytes) + "-byte scratch array");
    }
    q0 += qn;
  }
}

void Index::query(const float* queries, uint64_t nq, uint64_t d, const float* qlabels, uint64_t qcols,
                  float* out_dist, uint64_t* out_ids, hipStream_t s) {
  HDG_HIP(hipSetDevice(device_));
  const uint64_t G = check_query(queries, nq, d, qlabels, qcols, out_dist, out_ids);
  if (nq == 0) return;
  uint64_t chunk = std::max<uint64_t>(G, (cfg_.query_chunk / G) * G);
  last_w_ = w_;
  Work& w = *w_;
  w.st.reserve(nq * 6 * 8);
  w.st_chosen.reserve(nq * cfg_.sub_lists * 4);
  w.st_nq = nq;
  w.st_sub_lists = cfg_.sub_lists;
  w.ev_used = 0;
  if (!cfg_.graph) {
    for (uint64_t q0 = 0; q0 < nq; q0 += chunk) {
      uint64_t qn = std::min(chunk, nq - q0);
      query_chunk(queries, q0, qn, qlabels, qcols, G, out_dist, out_ids, s);
    }
    return;
  }
  // GRAPH 1: every chunk's launch sequence is captured once (after an uncaptured run that sizes the buffers)
  // and replayed on the internal stream, ordered after the caller's stream; one host synchronisation per
  // chunk at its end reads the abort flag (capacity overflow -> uncaptured rerun and recapture).
  if (!w.gstream) {
    HDG_HIP(hipStreamCreateWithFlags(&w.gstream, hipStreamNonBlocking));
    for (auto& e : w.gev) HDG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  w.abort_flag.reserve(4);
  const uint64_t digest = config_digest();
  if (w.cap_cfg != digest) {
    w.cap_cfg = digest;
    w.lcap = w.acap = 0;
    w.last_totL = w.last_totA = 0;
    w.need_mask = 0;
  }
  uint64_t ci = 0;
  for (uint64_t q0 = 0; q0 < nq; q0 += chunk, ++ci) {
    const uint64_t qn = std::min(chunk, nq - q0);
    if (w.graphs.size() <= ci) w.graphs.resize(ci + 1);
    Work::Captured& c = w.graphs[ci];
    const void* key[4] = {queries, qlabels, out_dist, out_ids};
    const bool valid = c.exec && std::equal(key, key + 4, c.key) && c.q0 == q0 && c.qn == qn && c.G == G &&
                       c.qcols == qcols && c.cfg == digest && c.gen == devbuf_generation();
    if (valid) {
      HDG_HIP(hipEventRecord(w.gev[0], s));
      HDG_HIP(hipStreamWaitEvent(w.gstream, w.gev[0], 0));
      HDG_HIP(hipGraphLaunch(c.exec, w.gstream));
      HDG_HIP(hipMemcpyAsync(&w.pinned[4], w.abort_flag.data(), 4, hipMemcpyDeviceToHost, w.gstream));
      HDG_HIP(hipMemcpyAsync(&w.pinned[5], w.err.data(), 4, hipMemcpyDeviceToHost, w.gstream));
      HDG_HIP(hipStreamSynchronize(w.gstream));
      if ((uint32_t)w.pinned[5] & 1u) fail(kInvalidArgument, "map filter label outside interval_map");
      if ((uint32_t)w.pinned[5] & 2u)
        fail(kUnsupported, "Morton split mode: a query box whose lower corner lies beyond the label range");
      if (((uint32_t)w.pinned[4] & 1u) == 0) continue;
    }
    // uncaptured run: results for this call, buffer sizes and the fallback kernels needed for the capture
    query_chunk(queries, q0, qn, qlabels, qcols, G, out_dist, out_ids, s);
    HDG_HIP(hipMemcpyAsync(&w.pinned[8], w.n_dense.data(), 5 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    w.lcap = std::max(w.lcap, w.last_totL + w.last_totL / 2 + 4096);
    // a chunk that needed no ADC is captured without the ADC stage (capacity 0: any ADC work aborts the replay)
    if (w.last_totA > 0) w.acap = std::max(w.acap, w.last_totA + w.last_totA / 2 + 4096);
    w.L.reserve(w.lcap * 4);
    w.Lkeys.reserve(w.lcap * 8);
    w.adist.reserve(w.acap * 4);
    w.akeys2.reserve(w.acap * 8);
    HDG_HIP(hipStreamSynchronize(s));
    {
      const uint32_t* nd = reinterpret_cast<const uint32_t*>(&w.pinned[8]);
      w.need_mask |= (nd[0] ? 1u : 0u) | (nd[1] ? 2u : 0u) | (nd[3] ? 4u : 0u) | (nd[4] ? 8u : 0u);
    }
    if (c.exec) {
      (void)hipGraphExecDestroy(c.exec);
      c.exec = nullptr;
    }
    // a buffer that grows while the capture records would leave earlier nodes on the freed block: such a
    // graph is dropped and the (now large enough) buffers are captured again
    for (int attempt = 0;; ++attempt) {
      const uint64_t gen0 = devbuf_generation();
      hipGraph_t g = nullptr;
      HDG_HIP(hipStreamBeginCapture(w.gstream, hipStreamCaptureModeThreadLocal));
      w.capturing = true;
      try {
        query_chunk(queries, q0, qn, qlabels, qcols, G, out_dist, out_ids, w.gstream);
      } catch (...) {
        w.capturing = false;
        (void)hipStreamEndCapture(w.gstream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      w.capturing = false;
      HDG_HIP(hipStreamEndCapture(w.gstream, &g));
      if (devbuf_generation() != gen0) {
        (void)hipGraphDestroy(g);
        if (attempt > 0) fail(kBadState, "query buffers kept growing during graph capture");
        continue;
      }
      const hipError_t ie = hipGraphInstantiate(&c.exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HDG_HIP(ie);
      break;
    }
    std::copy(key, key + 4, c.key);
    c.q0 = q0;
    c.qn = qn;
    c.G = G;
    c.qcols = qcols;
    c.cfg = digest;
    c.gen = devbuf_generation();
  }
}

uint64_t Index::config_digest() const {
  const uint64_t v[] = {cfg_.n_list, cfg_.sub_lists, cfg_.topk, cfg_.exps0, cfg_.precision, cfg_.probe_clusters,
                        cfg_.budget_model, cfg_.sparse_count, cfg_.sparse_max_tests, cfg_.adc_per, cfg_.adc_wgs,
                        cfg_.parity_batch, cfg_.query_chunk, cfg_.label_mask, cfg_.select_wave, cfg_.score_block,
                        cfg_.adc_split, cfg_.inbox_cap, cfg_.is_split, cfg_.n_split, cfg_.label_words,
                        graph_epoch_};
  uint64_t h = 1469598103934665603ull;
  for (uint64_t x : v) {
    h ^= x;
    h *= 1099511628211ull;
  }
  return h;
}

void Index::drop_graphs() {
  ++graph_epoch_;
  if (!w_) return;
  for (auto& c : w_->graphs)
    if (c.exec) (void)hipGraphExecDestroy(c.exec);
  w_->graphs.clear();
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
  if (d > kMaxDim) fail(kUnsupported, "vector dimension above " + std::to_string(kMaxDim));
  if (C > (uint64_t)dev::kSortCap) fail(kUnsupported, "clusters above 4096");
  if (cfg_.n_list > 1024) fail(kUnsupported, "n_list above 1024");
  if (topk > (uint64_t)dev::kSortCap || topk == 0) fail(kUnsupported, "topk must be in [1, 4096]");
  if (sub_lists == 0 || sub_lists > (uint32_t)dev::kSortCap) fail(kUnsupported, "sub_lists must be in [1, 4096]");
  if (g.n_valid < 1 || g.n_valid > 8) fail(kBadState, "no valid label dims");
  const uint32_t cap = (uint32_t)std::min<uint64_t>(CS, (uint64_t)n_list * S.max_subs_per_cluster);
  const uint64_t ng = cdiv(q0 + qn, G) - q0 / G;  // groups touched by this chunk
  const uint64_t g_base = q0 / G;
  const Dims D = make_dims(g);
  const bool sparse = cfg_.sparse_count != 0 && S.Nb > 0 && CS <= kSparseMaxCS;
  // PRECISION 1 refines a walked query from its in-box list, without Morton ranges or PQ (mode 3)
  const int inbox_mode = (cfg_.precision == 1 && sparse) ? 1 : 0;
  // ADC work split: items per thread (chunk = 256 * adc_per candidates) and persistent workgroups
  const uint32_t adc_per = cfg_.adc_per == 8 || cfg_.adc_per == 16 || cfg_.adc_per == 64 ? (uint32_t)cfg_.adc_per : 32u;
  const uint64_t adc_wgs = cfg_.adc_wgs ? cfg_.adc_wgs : 2048;

  // workspace
  w.ranges.reserve(qn * 2 * l * 4);
  w.boxes.reserve(qn * 2 * l * 4);
  w.box_lo.reserve((uint64_t)g.n_valid * qn * 4);
  w.box_w.reserve((uint64_t)g.n_valid * qn * 4);
  w.zlo.reserve(qn * 8);
  w.zhi.reserve(qn * 8);
  w.zlo_cnt.reserve(qn * 8);
  w.err.reserve(4);
  w.dense_list.reserve(qn * 4);
  w.n_dense.reserve(32);
  w.redo.reserve(qn * 4);
  w.counts.reserve(qn * CS * 4);
  if (cfg_.inbox_cap < 512 || cfg_.inbox_cap > (1u << 20)) fail(kUnsupported, "INBOX_CAP must be in [512, 2^20]");
  const uint32_t ibcap = (uint32_t)cfg_.inbox_cap;
  w.ib_pos.reserve(qn * ibcap * 4);
  if (!S.entries_packed) w.ib_sub.reserve(qn * ibcap * 4);  // packed entries list sub and position in one word
  w.n_ib.reserve(qn * 4);
  w.sub_list.reserve(qn * CS * 4);
  w.sub_cnt.reserve(qn * CS * 4);
  w.ns.reserve(qn * 4);
  w.list_c.reserve(qn * C * 4);
  w.cl_beg.reserve(qn * (C + 1) * 4);
  w.nnz_c.reserve(qn * 4);
  w.chosen_c.reserve(qn * n_list * 4);
  w.n_chosen_c.reserve(qn * 4);
  w.list_s.reserve(qn * cap * 4);
  w.list_scnt.reserve(qn * cap * 4);
  w.nnz_s.reserve(qn * 4);
  w.chosen_s.reserve(qn * sub_lists * 4);
  w.n_chosen_s.reserve(qn * 4);
  w.keys.reserve(qn * std::max<uint64_t>(C, cap) * 8);
  w.woff.reserve((qn + 1) * 8);
  w.gstat.reserve((ng + g_base + 1) * 8 * 8);
  // Morton split mode: each chosen sub-cluster's range becomes N_SPLIT pieces of the query's range table
  const uint32_t n_split = cfg_.is_split ? (uint32_t)cfg_.n_split : 0u;
  const uint32_t R = n_split ? sub_lists * n_split : sub_lists;
  w.rstart.reserve(qn * R * 4);
  w.rpre.reserve(qn * (R + 1) * 4);
  w.nrng.reserve(qn * 4);
  const uint32_t* nr_arg = n_split ? w.nrng.as<uint32_t>() : w.n_chosen_s.as<uint32_t>();
  w.cand.reserve(qn * 8);
  w.real.reserve(qn * 8);
  w.mode.reserve(qn * 4);
  const uint32_t sel_slots = (uint32_t)std::min<uint64_t>(qn, kSelSlots);
  w.sslot.reserve(qn * 4);
  w.smm.reserve(qn * 8);
  w.scnt.reserve(qn * 16);
  w.ghist.reserve((uint64_t)sel_slots * kSelBins * 4);
  w.kprime.reserve(qn * 8);
  w.lsize.reserve((qn + 1) * 8);
  w.asize.reserve((qn + 1) * 8);
  w.chunks.reserve((qn + 1) * 8);
  w.pad32.reserve(qn * 8);
  w.padmax.reserve(qn * 8);
  w.nL.reserve(qn * 4);
  w.bsum.reserve(3 * kScanBlocks * 8);
  if (cdiv(qn, 1024) > (uint64_t)kScanBlocks) fail(kUnsupported, "QUERY_CHUNK above 65536");
  uint32_t* abort_w = w.capturing ? w.abort_flag.as<uint32_t>() : nullptr;
  const uint32_t* abort_p = abort_w;
  // the chosen sub-cluster ids go straight into the call-wide array read by last_chosen_subs
  uint32_t* csid = w.st_chosen.as<uint32_t>() + q0 * sub_lists;
  auto scan_u64 = [&](uint64_t* a0, uint64_t* a1, uint64_t* a2, uint64_t n, const uint32_t* n32, const uint32_t* md,
                      uint32_t skip_mode, uint32_t* cap_abort = nullptr) {
    const unsigned nb = (unsigned)cdiv(n, 1024);
    hipLaunchKernelGGL(scan_local_kernel, dim3(nb), dim3(1024), 0, s, a0, a1, a2, n, n32, md, skip_mode,
                       w.bsum.as<unsigned long long>(), abort_p);
    hipLaunchKernelGGL(scan_fix_kernel, dim3(nb), dim3(1024), 0, s, a0, a1, a2, n, w.bsum.as<unsigned long long>(),
                       w.lcap, w.acap, cap_abort, abort_p);
  };
  unsigned long long* gstat = w.gstat.as<unsigned long long>() + g_base * 8;
  hipLaunchKernelGGL(clear_kernel, dim3((unsigned)cdiv(ng * 8, 256)), dim3(256), 0, s, gstat, ng * 8,
                     w.err.as<uint32_t>(), w.n_dense.as<uint32_t>(), abort_w);
  HDG_HIP(hipGetLastError());

  // dynamic LDS of the per-query select kernels: the sort buffer
  const uint32_t scap_c = sort_cap(n_list), scap_s = sort_cap(sub_lists), scap_f = sort_cap(topk);
  const unsigned dist_grid = 4096;
  if (!w.ev_init && !w.noalloc) {
    for (auto& e : w.ev) HDG_HIP(hipEventCreate(&e));
    w.ev_init = true;
  }
  int evn = 0;
  auto mark = [&]() {
    if (!w.capturing && w.ev_init) HDG_HIP(hipEventRecord(w.ev[evn++], s));
  };
  mark();
  // one thread per query, 64-thread blocks so that a 10k-query chunk spreads over more CUs
  hipLaunchKernelGGL(prep_kernel, dim3((unsigned)cdiv(qn, 64)), dim3(64), 0, s, qlabels, qcols, q0, qn, G, D,
                     f_type_.as<int>(), f_shift_.as<float>(), f_map_.as<float>(), f_map_len_, f_div_.as<int>(),
                     w.ranges.as<float>(), w.boxes.as<int>(), w.box_lo.as<uint32_t>(), w.box_w.as<uint32_t>(),
                     w.zlo.as<uint64_t>(), w.zhi.as<uint64_t>(), w.zlo_cnt.as<uint64_t>(), w.err.as<uint32_t>());
  HDG_HIP(hipGetLastError());
  mark();
  const bool mixed = cfg_.precision == 1;
  uint32_t* budget_T = nullptr;
  uint32_t* budget_S = nullptr;
  if (mixed && cfg_.budget_model == 1) {
    if (!budget_model_ready()) fail(kBadState, "BUDGET_MODEL = 1 but no budget model is loaded");
    w.budget_T.reserve(qn * 4);
    w.budget_S.reserve(qn * 4);
    budget_T = w.budget_T.as<uint32_t>();
    budget_S = w.budget_S.as<uint32_t>();
  }
  // PRECISION 1 with every non-zero cluster probed (PROBE_CLUSTERS and the budget model's cluster ladder at
  // least CLUSTERS): the chosen clusters' non-zero sub-clusters are all non-zero sub-clusters, so the walk's
  // ascending list is scored as it is (the order only breaks exact score ties) and the cluster choice and
  // expansion are skipped.
  const uint32_t kc = cfg_.probe_clusters > 0 ? std::min<uint32_t>(n_list, (uint32_t)cfg_.probe_clusters) : n_list;
  const bool all_clusters = mixed && kc >= C && (!budget_T || budget_probes_all((uint32_t)C));
  // Graph capture leaves out the fallback kernels the sizing runs never needed (bits: 1 dense count, 2 workgroup
  // walk, 4 workgroup scored selection, 8 workgroup final pick); a kernel that would hand such a kernel a query
  // sets the abort flag instead, every later kernel returns, and the chunk is rerun uncaptured.  Bits 1, 2 and 4
  // need every kernel between the walk and the capacity check to honour the flag: the PRECISION 1 path with
  // in-box refinement, wave selection and every cluster probed.
  const bool lean_ok = mixed && inbox_mode && cfg_.select_wave && all_clusters;
  const uint32_t skip = w.capturing && cfg_.select_wave ? (~w.need_mask & (lean_ok ? 0xFu : 0x8u)) : 0u;
  // PRECISION 1 scores queries x (sub-)centers in blocks of SCORE_BLOCK queries through one score buffer, small
  // enough to stay in the memory-side cache between the GEMM that writes it and the selection that reads it
  const uint64_t score_block = std::min<uint64_t>(qn, std::max<uint64_t>(cfg_.score_block, 128));
  // PRECISION 1: per query, bit words of its non-zero sub-clusters and their ranks (compact score rows)
  const uint32_t nws = (uint32_t)((CS + 31) / 32);
  uint32_t* nzb = nullptr;
  uint32_t* nzp = nullptr;
  if (mixed) {
    w.nzb.reserve(qn * nws * 4);
    w.nzp.reserve(qn * nws * 4);
    nzb = w.nzb.as<uint32_t>();
    nzp = w.nzp.as<uint32_t>();
    w.S.reserve(score_block * std::max<uint64_t>(C, CS) * 4);
    w.feat.reserve(qn * kFeatDists * 4);
    w.qb16.reserve(qn * S.Dp * 2);
    launch_to_bf16(queries, q0, qn, d, S.Dp, w.qb16.as<uint16_t>(), s);
    HDG_HIP(hipGetLastError());
  }

  // ---- A3/A4: non-zero sub-clusters and clusters of every query ----
  {
    Buckets B;
    for (int j = 0; j < kMaxDims; ++j) {
      B.M[j] = S.Mv[j];
      B.nb[j] = S.nb[j];
      B.stride[j] = S.bstride[j];
    }
    B.cells16 = S.cells16 ? 1u : 0u;
    B.packed = S.entries_packed ? 1u : 0u;
    B.pk_shift = S.pk_shift;
    const uint32_t nw = (uint32_t)((CS + 31) / 32);
    const size_t lds = (CS + 2 * nw) * sizeof(uint32_t);
    const uint32_t max_tests = (uint32_t)std::min<uint64_t>(cfg_.sparse_max_tests, 0xFFFFFFF0ull);
    const unsigned grid = (unsigned)std::min<uint64_t>(qn, 2048);
    const unsigned wgrid = (unsigned)std::min<uint64_t>(cdiv(qn, 4), 2048);
    // the wave walk's hand-overs: few on most workloads, nearly every query when boxes hold thousands of
    // points; workgroups without a hand-over exit at once, so the grid is sized for the many
    const unsigned rgrid = grid;
    const uint32_t* cells = S.cells.as<uint32_t>();
    dim3 dgrid((unsigned)cdiv(qn, 256), (unsigned)cdiv(CS, kCountSegChunk));
    CoarseSums sums;
    if (S.sat_per) {
      sums.t = S.sat.as<uint32_t>();
      sums.per = S.sat_per;
      for (int j = 0; j < g.n_valid; ++j) {
        sums.sg[j] = S.sg[j];
        sums.Mv[j] = S.Mv[j];
        sums.stride[j] = S.sstride[j];
      }
    }
    switch (g.n_valid) {
#define HDG_A3_CASE(NV)                                                                                            \
  case NV:                                                                                                         \
    if (sparse) {                                                                                                  \
      if (cfg_.select_wave)                                                                                        \
        hipLaunchKernelGGL(sparse_count_wave_kernel<NV>, dim3(wgrid), dim3(256), 0, s, B, S.bstart.as<uint32_t>(), \
                           S.ecell.as<uint32_t>(), S.esub.as<uint32_t>(), S.N, S.sub_off.as<uint32_t>(),           \
                           (uint32_t)CS, (uint32_t)C, w.box_lo.as<uint32_t>(), w.box_w.as<uint32_t>(), qn, G, q0,  \
                           max_tests, S.epos.as<uint32_t>(), ibcap, w.ib_pos.as<uint32_t>(),                   \
                           w.ib_sub.as<uint32_t>(), w.n_ib.as<uint32_t>(), w.sub_list.as<uint32_t>(),              \
                           w.sub_cnt.as<uint32_t>(), w.ns.as<uint32_t>(), w.list_c.as<uint32_t>(),                 \
                           w.cl_beg.as<uint32_t>(), w.nnz_c.as<uint32_t>(), w.dense_list.as<uint32_t>(),           \
                           w.n_dense.as<uint32_t>(), w.gstat.as<unsigned long long>(), nzb, nzp,                   \
                           w.redo.as<uint32_t>(), skip & 3u, abort_w);                                             \
      if (!cfg_.select_wave || !(skip & 2u))                                                                       \
      hipLaunchKernelGGL(sparse_count_kernel<NV>, dim3(cfg_.select_wave ? rgrid : grid), dim3(256), lds, s, B,      \
                         S.bstart.as<uint32_t>(),                                                                  \
                         S.ecell.as<uint32_t>(), S.esub.as<uint32_t>(), S.N, S.sub_off.as<uint32_t>(),             \
                         (uint32_t)CS, (uint32_t)C, w.box_lo.as<uint32_t>(), w.box_w.as<uint32_t>(), qn, G, q0,    \
                         max_tests, S.epos.as<uint32_t>(), ibcap, w.ib_pos.as<uint32_t>(),           \
                         w.ib_sub.as<uint32_t>(), w.n_ib.as<uint32_t>(), w.sub_list.as<uint32_t>(), w.sub_cnt.as<uint32_t>(), w.ns.as<uint32_t>(),                 \
                         w.list_c.as<uint32_t>(), w.cl_beg.as<uint32_t>(), w.nnz_c.as<uint32_t>(),                 \
                         w.dense_list.as<uint32_t>(), w.n_dense.as<uint32_t>(), w.gstat.as<unsigned long long>(), \
                         nzb, nzp, cfg_.select_wave ? w.redo.as<uint32_t>() : nullptr,                            \
                         cfg_.select_wave ? w.n_dense.as<uint32_t>() + 1 : nullptr, abort_p);                     \
    } else {                                                                                                       \
      hipLaunchKernelGGL(iota_kernel, dim3((unsigned)cdiv(qn, 256)), dim3(256), 0, s, qn,                          \
                         w.dense_list.as<uint32_t>(), w.n_dense.as<uint32_t>(), w.n_ib.as<uint32_t>());           \
    }                                                                                                              \
    if (!(skip & 1u))                                                                                              \
    hipLaunchKernelGGL(dense_count_kernel<NV>, dgrid, dim3(256), 0, s, cells, S.N, S.zcode.as<uint64_t>(),          \
                       S.seg_off.as<uint32_t>(), CS, w.box_lo.as<uint32_t>(), w.box_w.as<uint32_t>(),              \
                       w.zlo_cnt.as<uint64_t>(), w.zhi.as<uint64_t>(), qn, w.dense_list.as<uint32_t>(),            \
                       w.n_dense.as<uint32_t>(), w.counts.as<uint32_t>(), abort_p, sums);                          \
    break;
      HDG_A3_CASE(1)
      HDG_A3_CASE(2)
      HDG_A3_CASE(3)
      HDG_A3_CASE(4)
      HDG_A3_CASE(5)
      HDG_A3_CASE(6)
      HDG_A3_CASE(7)
      HDG_A3_CASE(8)
#undef HDG_A3_CASE
      default:
        fail(kBadState, "no valid label dims");
    }
    HDG_HIP(hipGetLastError());
    if (!(skip & 1u))
    hipLaunchKernelGGL(dense_emit_kernel, dim3((unsigned)qn), dim3(256), 0, s, w.counts.as<uint32_t>(), (uint32_t)CS,
                       S.sub_off.as<uint32_t>(), (uint32_t)C, G, q0, w.dense_list.as<uint32_t>(),
                       w.n_dense.as<uint32_t>(), w.sub_list.as<uint32_t>(), w.sub_cnt.as<uint32_t>(),
                       w.ns.as<uint32_t>(), w.list_c.as<uint32_t>(), w.cl_beg.as<uint32_t>(), w.nnz_c.as<uint32_t>(),
                       w.gstat.as<unsigned long long>(), nzb, nzp, abort_p);
    HDG_HIP(hipGetLastError());
  }
  mark();
  // learned budget (PRECISION 1): the cluster scoring has written the features (the 32 smallest coarse scores
  // and |q|^2); after the sub-cluster selection the model trims each query's sorted choice to its budget
  auto budgets = [&]() {
    capture_features(q0, qn, l, w.ranges.as<float>(), w.box_w.as<uint32_t>(), g.n_valid, w.feat.as<float>(), s);
    if (budget_T)
      run_budget_model(qn, l, w.ranges.as<float>(), w.box_lo.as<uint32_t>(), w.box_w.as<uint32_t>(), g.n_valid,
                       w.feat.as<float>(), w.n_chosen_s.as<uint32_t>(), budget_T, budget_S, s, abort_p);
  };
  const uint32_t* Ls = all_clusters ? w.sub_list.as<uint32_t>() : w.list_s.as<uint32_t>();
  const uint32_t* Lcnt = all_clusters ? w.sub_cnt.as<uint32_t>() : w.list_scnt.as<uint32_t>();
  const uint32_t* Lnnz = all_clusters ? w.ns.as<uint32_t>() : w.nnz_s.as<uint32_t>();
  // learned budget (PRECISION 1): features of each query -- filter widths, grid-cell extents and the 32 smallest
  // BF16 coarse scores among its non-zero clusters (+ |q|^2), written by the cluster selection
  const bool need_feat = mixed && (budget_T || budget_capturing());
  // ---- A5: cluster choice ----
  if (!mixed) {
    scan_u64(w.woff.as<uint64_t>(), nullptr, nullptr, qn, w.nnz_c.as<uint32_t>(), nullptr, 0u);
    launch_pair_dist<0>(dist_grid, s, queries, q0, d, S.centers.as<float>(), w.woff.as<uint64_t>(), qn,
                        w.list_c.as<uint32_t>(), C, nullptr, w.nnz_c.as<uint32_t>(), nullptr, w.keys.as<uint64_t>(),
                        nullptr);
    hipLaunchKernelGGL(coarse_pick_kernel, dim3((unsigned)qn), dim3(256), scap_c * 8, s, (uint32_t)C, n_list, G, q0,
                       scap_c,
                       w.nnz_c.as<uint32_t>(), w.gstat.as<unsigned long long>(), w.keys.as<uint64_t>(),
                       w.chosen_c.as<uint32_t>(), w.n_chosen_c.as<uint32_t>());
  } else if (!all_clusters) {
    // the coarse scores pick the clusters
    for (uint64_t b0 = 0; b0 < qn; b0 += score_block) {
      const uint64_t bn = std::min(score_block, qn - b0);
      launch_coarse_gemm(w.qb16.as<uint16_t>() + b0 * S.Dp, bn, S.Dp, S.centers_bf16.as<uint16_t>(),
                         S.center_norm.as<float>(), (uint32_t)C, w.S.as<float>(), s, nullptr, nullptr, 0, abort_p);
      launch_scored_select(queries, q0 + b0, bn, d, w.list_c.as<uint32_t>() + b0 * C, (uint32_t)C,
                           w.nnz_c.as<uint32_t>() + b0, w.S.as<float>(), C, kc, nullptr, n_list,
                           w.chosen_c.as<uint32_t>() + b0 * n_list, w.n_chosen_c.as<uint32_t>() + b0,
                           need_feat ? w.feat.as<float>() + b0 * kFeatDists : nullptr,
                           w.keys.as<uint64_t>() + b0 * C, cfg_.select_wave != 0, s, nullptr, nullptr, 0, abort_w,
                           w.n_dense.as<uint32_t>() + 3, (skip & 4u) != 0);
    }
  } else if (need_feat) {
    // every non-zero cluster is probed, so nothing is chosen here, but the budget features still come from the
    // coarse scores of the query's non-zero clusters
    for (uint64_t b0 = 0; b0 < qn; b0 += score_block) {
      const uint64_t bn = std::min(score_block, qn - b0);
      launch_coarse_gemm(w.qb16.as<uint16_t>() + b0 * S.Dp, bn, S.Dp, S.centers_bf16.as<uint16_t>(),
                         S.center_norm.as<float>(), (uint32_t)C, w.S.as<float>(), s, nullptr, nullptr, 0, abort_p);
      if (C <= 512) {  // one wave per query, 32 smallest only
        launch_coarse_features(queries, q0 + b0, bn, d, w.list_c.as<uint32_t>() + b0 * C, (uint32_t)C,
                               w.nnz_c.as<uint32_t>() + b0, w.S.as<float>(), C, w.feat.as<float>() + b0 * kFeatDists,
                               s, abort_w);
        continue;
      }
      launch_scored_select(queries, q0 + b0, bn, d, w.list_c.as<uint32_t>() + b0 * C, (uint32_t)C,
                           w.nnz_c.as<uint32_t>() + b0, w.S.as<float>(), C, 1, nullptr, n_list,
                           w.chosen_c.as<uint32_t>() + b0 * n_list, w.n_chosen_c.as<uint32_t>() + b0,
                           w.feat.as<float>() + b0 * kFeatDists, w.keys.as<uint64_t>() + b0 * C,
                           cfg_.select_wave != 0, s, nullptr, nullptr, 0, abort_w, w.n_dense.as<uint32_t>() + 3,
                           (skip & 4u) != 0);
    }
  }
  HDG_HIP(hipGetLastError());
  mark();
  // ---- A6: sub-cluster choice ----
  if (!all_clusters)
    hipLaunchKernelGGL(sub_expand_kernel, dim3((unsigned)qn), dim3(256), 0, s, (uint32_t)CS, (uint32_t)C,
                       S.sub_off.as<uint32_t>(), n_list, G, q0, w.list_c.as<uint32_t>(), w.cl_beg.as<uint32_t>(),
                       w.sub_list.as<uint32_t>(), w.sub_cnt.as<uint32_t>(), w.chosen_c.as<uint32_t>(),
                       w.n_chosen_c.as<uint32_t>(), cap, w.list_s.as<uint32_t>(), w.list_scnt.as<uint32_t>(),
                       w.nnz_s.as<uint32_t>(), w.gstat.as<unsigned long long>());
  HDG_HIP(hipGetLastError());
  mark();
  if (!mixed) {
    scan_u64(w.woff.as<uint64_t>(), nullptr, nullptr, qn, w.nnz_s.as<uint32_t>(), nullptr, 0u);
    launch_pair_dist<0>(dist_grid, s, queries, q0, d, S.sub_centers.as<float>(), w.woff.as<uint64_t>(), qn,
                        w.list_s.as<uint32_t>(), (uint64_t)cap, nullptr, w.nnz_s.as<uint32_t>(), nullptr,
                        w.keys.as<uint64_t>(), nullptr);
    hipLaunchKernelGGL(sub_pick_kernel, dim3((unsigned)qn), dim3(256), scap_s * 8, s, sub_lists, G, q0, cap, scap_s,
                       w.nnz_s.as<uint32_t>(), w.gstat.as<unsigned long long>(), budget_S, w.keys.as<uint64_t>(),
                       w.chosen_s.as<uint32_t>(), w.n_chosen_s.as<uint32_t>());
  } else {
    for (uint64_t b0 = 0; b0 < qn; b0 += score_block) {
      const uint64_t bn = std::min(score_block, qn - b0);
      launch_coarse_gemm(w.qb16.as<uint16_t>() + b0 * S.Dp, bn, S.Dp, S.sub_centers_bf16.as<uint16_t>(),
                         S.sub_center_norm.as<float>(), (uint32_t)CS, w.S.as<float>(), s, nzb + b0 * nws,
                         nzp + b0 * nws, nws, abort_p);
      launch_scored_select(queries, q0 + b0, bn, d, Ls + b0 * cap, cap, Lnnz + b0, w.S.as<float>(), CS, sub_lists,
                           nullptr, sub_lists, w.chosen_s.as<uint32_t>() + b0 * sub_lists,
                           w.n_chosen_s.as<uint32_t>() + b0, nullptr,
                           w.keys.as<uint64_t>() + b0 * cap, cfg_.select_wave != 0, s, nzb + b0 * nws,
                           nzp + b0 * nws, nws, abort_w, w.n_dense.as<uint32_t>() + 3, (skip & 4u) != 0,
                           all_clusters);
    }
    budgets();
  }
  HDG_HIP(hipGetLastError());
  mark();
  if (inbox_mode && cfg_.select_wave)
    hipLaunchKernelGGL(ranges_inbox_wave_kernel, dim3((unsigned)cdiv(qn, 4)), dim3(256), 0, s, qn, cap, Ls, Lcnt,
                       sub_lists, w.chosen_s.as<uint32_t>(), w.n_chosen_s.as<uint32_t>(), csid,
                       w.rpre.as<uint32_t>(), w.cand.as<uint64_t>(), w.real.as<uint64_t>(), w.n_ib.as<uint32_t>(),
                       abort_p);
  // without dense-counted or workgroup-walked queries every query was refined from its in-box list by the
  // wave kernels: the Morton-range kernels below have nothing to do
  const bool all_walked = (skip & 3u) == 3u;
  if (!all_walked)
  hipLaunchKernelGGL(ranges_kernel, dim3((unsigned)qn), dim3(256), 0, s, S.zcode.as<uint64_t>(),
                     S.seg_off.as<uint32_t>(), S.cells.as<uint32_t>(), S.N, g.n_valid, w.box_lo.as<uint32_t>(),
                     w.box_w.as<uint32_t>(), qn, w.zlo_cnt.as<uint64_t>(), cap, Ls, Lcnt, sub_lists, G,
                     q0, w.chosen_s.as<uint32_t>(), w.n_chosen_s.as<uint32_t>(), w.zlo.as<uint64_t>(),
                     w.zhi.as<uint64_t>(), csid, w.rstart.as<uint32_t>(),
                     w.rpre.as<uint32_t>(), w.cand.as<uint64_t>(), w.real.as<uint64_t>(),
                     w.gstat.as<unsigned long long>(), w.n_ib.as<uint32_t>(),
                     inbox_mode && cfg_.select_wave ? 2 : inbox_mode, n_split ? 1 : 0, abort_p);
  HDG_HIP(hipGetLastError());
  if (n_split && !all_walked) {
    hipLaunchKernelGGL(split_ranges_kernel, dim3((unsigned)qn), dim3(256), 0, s, S.zcode.as<uint64_t>(),
                       S.seg_off.as<uint32_t>(), S.cells.as<uint32_t>(), S.N, g.n_valid, g.max_bits,
                       w.box_lo.as<uint32_t>(), w.box_w.as<uint32_t>(), qn, w.zlo.as<uint64_t>(),
                       w.zhi.as<uint64_t>(), w.zlo_cnt.as<uint64_t>(), sub_lists, n_split, csid,
                       w.n_chosen_s.as<uint32_t>(), w.n_ib.as<uint32_t>(), inbox_mode, w.rstart.as<uint32_t>(),
                       w.rpre.as<uint32_t>(), w.nrng.as<uint32_t>(), w.cand.as<uint64_t>(),
                       w.gstat.as<unsigned long long>(), G, q0, w.err.as<uint32_t>(), abort_p);
    HDG_HIP(hipGetLastError());
  }
  mark();
  hipLaunchKernelGGL(plan_kernel, dim3((unsigned)cdiv(qn, 64)), dim3(64), 0, s, qn, q0, G, l, cfg_.exps0,
                     topk, w.cand.as<uint64_t>(), w.gstat.as<unsigned long long>(), w.mode.as<uint32_t>(),
                     w.kprime.as<uint64_t>(), w.lsize.as<uint64_t>(), w.asize.as<uint64_t>(),
                     w.chunks.as<uint64_t>(), w.pad32.as<uint64_t>(), w.padmax.as<uint64_t>(), w.n_ib.as<uint32_t>(),
                     inbox_mode, adc_per * (uint32_t)kWG, w.sslot.as<uint32_t>(), w.smm.as<uint32_t>(),
                     w.scnt.as<uint32_t>(), w.n_dense.as<uint32_t>() + 2, sel_slots, cfg_.adc_split, abort_p);
  HDG_HIP(hipGetLastError());
  // captured: fixed capacities; a chunk that needs more raises the abort flag in the scan's fix-up
  scan_u64(w.lsize.as<uint64_t>(), w.asize.as<uint64_t>(), w.chunks.as<uint64_t>(), qn, nullptr, nullptr, 0, abort_w);
  HDG_HIP(hipGetLastError());
  uint64_t totL, totA, totChunks;
  if (!w.capturing) {
    // the one host synchronisation of a chunk: sizes of the data-dependent buffers
    HDG_HIP(hipMemcpyAsync(&w.pinned[0], w.lsize.as<uint64_t>() + qn, 8, hipMemcpyDeviceToHost, s));
    HDG_HIP(hipMemcpyAsync(&w.pinned[1], w.asize.as<uint64_t>() + qn, 8, hipMemcpyDeviceToHost, s));
    HDG_HIP(hipMemcpyAsync(&w.pinned[2], w.chunks.as<uint64_t>() + qn, 8, hipMemcpyDeviceToHost, s));
    HDG_HIP(hipMemcpyAsync(&w.pinned[3], w.err.data(), 4, hipMemcpyDeviceToHost, s));
    HDG_HIP(hipStreamSynchronize(s));
    mark();
    if ((uint32_t)w.pinned[3] & 1u) fail(kInvalidArgument, "map filter label outside interval_map");
    if ((uint32_t)w.pinned[3] & 2u)
      fail(kUnsupported, "Morton split mode: a query box whose lower corner lies beyond the label range");
    totL = w.pinned[0];
    totA = w.pinned[1];
    totChunks = w.pinned[2];
    w.last_totL = std::max(w.last_totL, totL);
    w.last_totA = std::max(w.last_totA, totA);
  } else {
    totL = w.lcap;
    totA = w.acap;
    totChunks = w.acap ? 1 : 0;  // the persistent ADC grid reads the real count on the device
  }
  w.L.reserve(std::max<uint64_t>(totL, 1) * 4);
  w.Lkeys.reserve(std::max<uint64_t>(totL, 1) * 8);
  w.adist.reserve(std::max<uint64_t>(totA, 1) * 4);
  w.akeys2.reserve(std::max<uint64_t>(totA, 1) * 8);
  // split mode keeps the keep-all candidates its pieces give (no in-box replacement)
  const int integral = n_split ? -1 : (S.labels_integral ? 1 : 0);
  const uint32_t* ib_sub_arg = S.entries_packed ? nullptr : w.ib_sub.as<uint32_t>();
  LabelWords lw;
  lw.w = S.lab_words && cfg_.label_words ? S.labw.as<uint64_t>() : nullptr;
  lw.lay = S.lab_words ? S.lw_dev.as<LabelWordLayout>() : nullptr;
  lw.guard = S.lw.guard;
  if (sparse && cfg_.select_wave) {
    auto kern = S.entries_packed ? (lw.w ? inbox_keep_wave_kernel<true, true> : inbox_keep_wave_kernel<true, false>)
                                 : (lw.w ? inbox_keep_wave_kernel<false, true> : inbox_keep_wave_kernel<false, false>);
    hipLaunchKernelGGL(kern, dim3((unsigned)cdiv(qn, kKeepWaves)), dim3(64 * kKeepWaves),
                       S.entries_packed ? CS * 4 : 0, s, qn, l, integral,
                       (uint32_t)CS, sub_lists, w.mode.as<uint32_t>(), w.n_ib.as<uint32_t>(), w.zlo.as<uint64_t>(),
                       w.zlo_cnt.as<uint64_t>(), w.n_chosen_s.as<uint32_t>(), csid, ibcap,
                       w.ib_pos.as<uint32_t>(), ib_sub_arg, S.seg_off.as<uint32_t>(), S.pk_shift, S.lab.as<float>(),
                       lw, w.ranges.as<float>(), w.lsize.as<uint64_t>(), w.L.as<uint32_t>(), w.nL.as<uint32_t>(), abort_p);
  }
  else if (sparse)
    hipLaunchKernelGGL(inbox_keep_kernel, dim3((unsigned)qn), dim3(256), 0, s, l, integral, (uint32_t)CS, sub_lists,
                       w.mode.as<uint32_t>(), w.n_ib.as<uint32_t>(), w.zlo.as<uint64_t>(), w.zlo_cnt.as<uint64_t>(),
                       w.n_chosen_s.as<uint32_t>(), csid, ibcap, w.ib_pos.as<uint32_t>(), ib_sub_arg,
                       S.seg_off.as<uint32_t>(), S.pk_shift, S.lab.as<float>(), lw, w.ranges.as<float>(),
                       w.lsize.as<uint64_t>(), w.L.as<uint32_t>(), w.nL.as<uint32_t>(), abort_p);
  if (!all_walked)
  hipLaunchKernelGGL(keep_all_kernel, dim3((unsigned)qn), dim3(256), (2 * R + 1) * 4, s, qn, l, integral,
                     R, w.mode.as<uint32_t>(), w.n_ib.as<uint32_t>(), w.zlo.as<uint64_t>(),
                     w.zlo_cnt.as<uint64_t>(), nr_arg, w.rstart.as<uint32_t>(),
                     w.rpre.as<uint32_t>(), S.lab.as<float>(), lw, w.ranges.as<float>(), w.lsize.as<uint64_t>(),
                     w.L.as<uint32_t>(), w.nL.as<uint32_t>(), abort_p);
  HDG_HIP(hipGetLastError());
  mark();
  if (totChunks > 0) {
    // LDS: the LUT (one phase, at most kLutPhaseBytes) and the range table
    const uint64_t lut_entry = mixed ? 2 : 4;
    const uint32_t lut_bytes =
        (uint32_t)((std::min<uint64_t>(kLutPhaseBytes, S.M * S.K * lut_entry) + 15) & ~15ull);
    const size_t adc_lds = lut_bytes + (R <= kAdcLdsRanges ? (2 * (size_t)R + 1) * 4 : 0);
    const unsigned adc_grid = (unsigned)(w.capturing ? adc_wgs : std::min<uint64_t>(totChunks, adc_wgs));
    // the streaming variant needs the whole LUT in one LDS phase and 32-code steps
    uint64_t mp_h = kLutPhaseBytes / (S.K * lut_entry);
    if (mp_h > 4) mp_h &= ~3ull;
    // streaming variants: the whole LUT in one phase; 32-code steps (16-byte loads) or, for code rows of at
    // most 64 codes with pq_dim % 8 == 0, 8-code steps (8-byte loads)
    int stream = mp_h < S.M || (mp_h % 16) != 0 ? 0
                 : (S.M % 32) == 0                ? 32
                 : ((S.M % 8) == 0 && S.M <= 24) ? 8
                                                  : 0;
    // the streaming variants read the 64-row blocked codes
    if (S.codes_wb != 16 || !S.codes_blk.data()) stream = 0;
    const uint8_t* codes_arg = stream ? S.codes_blk.as<uint8_t>() : S.codes.as<uint8_t>();
    // the persistent grid is at most what is resident at once, so every workgroup's share starts at launch
#define HDG_ADC1(LT, PER, KC, ST)                                                                                 \
  hipLaunchKernelGGL((adc_scan_kernel<LT, PER, KC, ST>),                                                          \
                     dim3(resident_grid(reinterpret_cast<const void*>(&adc_scan_kernel<LT, PER, KC, ST>), adc_lds,  \
                                        adc_grid)),                                                               \
                     dim3(256), adc_lds, s, qn, q0, queries, d,                                                    \
                     S.M, S.K, S.pq_len, S.pq_centers.as<float>(), codes_arg, R,                                      \
                     w.mode.as<uint32_t>(), nr_arg, w.rstart.as<uint32_t>(),                                          \
                     w.rpre.as<uint32_t>(), w.cand.as<uint64_t>(), w.L.as<uint32_t>(), w.lsize.as<uint64_t>(),       \
                     w.nL.as<uint32_t>(), w.chunks.as<uint64_t>(), w.asize.as<uint64_t>(), w.adist.as<uint32_t>(),   \
                     lut_bytes, w.smm.as<uint32_t>(), abort_p)
#define HDG_ADC(LT, PER)                                                                                          \
  do {                                                                                                            \
    if (S.K == 256 && stream == 32) HDG_ADC1(LT, PER, 256, 32);                                                   \
    else if (S.K == 256 && stream == 8) HDG_ADC1(LT, PER, 256, 8);                                                \
    else if (S.K == 256) HDG_ADC1(LT, PER, 256, 0);                                                               \
    else if (stream == 32) HDG_ADC1(LT, PER, 0, 32);                                                              \
    else if (stream == 8) HDG_ADC1(LT, PER, 0, 8);                                                                \
    else HDG_ADC1(LT, PER, 0, 0);                                                                                 \
  } while (0)
    switch (adc_per) {
      case 8:
        if (mixed) HDG_ADC(__half, 8); else HDG_ADC(float, 8);
        break;
      case 16:
        if (mixed) HDG_ADC(__half, 16); else HDG_ADC(float, 16);
        break;
      case 64:
        if (mixed) HDG_ADC(__half, 64); else HDG_ADC(float, 64);
        break;
      default:
        if (mixed) HDG_ADC(__half, 32); else HDG_ADC(float, 32);
        break;
    }
#undef HDG_ADC
#undef HDG_ADC1
    HDG_HIP(hipGetLastError());
  }
  mark();
  if (totChunks > 0) {
    const unsigned sel_grid = (unsigned)std::max<uint64_t>(1, cdiv(totA, kSelTile));
    hipLaunchKernelGGL(sel_zero_kernel, dim3(sel_slots * (kSelBins / 4) / 256), dim3(256), 0, s,
                       w.n_dense.as<uint32_t>() + 2, sel_slots, w.ghist.as<uint32_t>());
    hipLaunchKernelGGL(sel_hist_kernel, dim3(sel_grid), dim3(256), 0, s, qn, w.asize.as<uint64_t>(),
                       w.sslot.as<uint32_t>(), w.smm.as<uint32_t>(), w.adist.as<uint32_t>(), w.ghist.as<uint32_t>(),
                       abort_p);
    hipLaunchKernelGGL(lw.w ? sel_split_kernel<true> : sel_split_kernel<false>, dim3(sel_grid), dim3(256), 0, s, qn, l, R,
                       w.asize.as<uint64_t>(),
                       w.sslot.as<uint32_t>(), w.smm.as<uint32_t>(), w.kprime.as<uint64_t>(), w.adist.as<uint32_t>(),
                       w.ghist.as<uint32_t>(), nr_arg, w.rstart.as<uint32_t>(),
                       w.rpre.as<uint32_t>(), S.lab.as<float>(), lw, w.ranges.as<float>(), w.lsize.as<uint64_t>(),
                       w.L.as<uint32_t>(), w.akeys2.as<uint64_t>(), w.scnt.as<uint32_t>(), abort_p);
    HDG_HIP(hipGetLastError());
    hipLaunchKernelGGL(adc_select_kernel, dim3((unsigned)qn), dim3(256), 0, s, qn, l, R,
                       w.mode.as<uint32_t>(), w.kprime.as<uint64_t>(), w.cand.as<uint64_t>(),
                       nr_arg, w.rstart.as<uint32_t>(), w.rpre.as<uint32_t>(), S.lab.as<float>(), lw,
                       w.ranges.as<float>(), w.asize.as<uint64_t>(), w.Lkeys.as<uint64_t>(), w.adist.as<uint32_t>(),
                       w.akeys2.as<uint64_t>(), w.lsize.as<uint64_t>(), w.L.as<uint32_t>(), w.nL.as<uint32_t>(),
                       w.sslot.as<uint32_t>(), w.smm.as<uint32_t>(), w.scnt.as<uint32_t>(), abort_p);
    HDG_HIP(hipGetLastError());
  }
  mark();
  // ---- A13: exact re-rank of the refine lists, final top-k ----
  scan_u64(w.woff.as<uint64_t>(), nullptr, nullptr, qn, w.nL.as<uint32_t>(), w.mode.as<uint32_t>(), 2u);
  mark();
  if (d >= kStreamRowsMinDim)
    launch_pair_dist<1, true>(dist_grid, s, queries, q0, d, S.vec.as<float>(), w.woff.as<uint64_t>(), qn,
                              w.L.as<uint32_t>(), (uint64_t)0, w.lsize.as<uint64_t>(), w.nL.as<uint32_t>(),
                              S.zid.as<uint32_t>(), w.Lkeys.as<uint64_t>(), abort_p);
  else
    launch_pair_dist<1, false>(dist_grid, s, queries, q0, d, S.vec.as<float>(), w.woff.as<uint64_t>(), qn,
                               w.L.as<uint32_t>(), (uint64_t)0, w.lsize.as<uint64_t>(), w.nL.as<uint32_t>(),
                               S.zid.as<uint32_t>(), w.Lkeys.as<uint64_t>(), abort_p);
  mark();
  int64_t pick_skip = -1;
  if (cfg_.select_wave) {
    hipLaunchKernelGGL(final_pick_wave_kernel, dim3((unsigned)cdiv(qn, 4)), dim3(256), 0, s, q0, qn,
                       S.zid.as<uint32_t>(), l, topk, w.mode.as<uint32_t>(), w.kprime.as<uint64_t>(),
                       w.lsize.as<uint64_t>(), w.L.as<uint32_t>(), w.nL.as<uint32_t>(), w.pad32.as<uint64_t>(),
                       w.padmax.as<uint64_t>(), w.Lkeys.as<uint64_t>(), w.adist.as<uint32_t>(),
                       w.asize.as<uint64_t>(), out_dist, out_ids, abort_p, w.n_dense.as<uint32_t>() + 4,
                       (skip & 8u) ? abort_w : nullptr, w.cand.as<uint64_t>(), w.real.as<uint64_t>(),
                       w.nnz_c.as<uint32_t>(), Lnnz, w.st.as<uint64_t>());
    pick_skip = kWavePickMax;
  }
  if (!(skip & 8u))
  hipLaunchKernelGGL(final_pick_kernel, dim3((unsigned)qn), dim3(256), scap_f * 8, s, q0, scap_f, S.zid.as<uint32_t>(), l,
                     topk,
                     w.mode.as<uint32_t>(), w.kprime.as<uint64_t>(), w.lsize.as<uint64_t>(), w.L.as<uint32_t>(),
                     w.nL.as<uint32_t>(), w.pad32.as<uint64_t>(), w.padmax.as<uint64_t>(), w.Lkeys.as<uint64_t>(),
                     w.adist.as<uint32_t>(), w.asize.as<uint64_t>(), out_dist, out_ids, abort_p, pick_skip);
  HDG_HIP(hipGetLastError());
  mark();
  if (!cfg_.select_wave)  // the wave final pick writes the counters itself
  hipLaunchKernelGGL(stats_kernel, dim3((unsigned)cdiv(qn, 256)), dim3(256), 0, s, qn, q0, w.cand.as<uint64_t>(),
                     w.real.as<uint64_t>(), w.nL.as<uint32_t>(), w.nnz_c.as<uint32_t>(), Lnnz,
                     w.mode.as<uint32_t>(), w.kprime.as<uint64_t>(), w.st.as<uint64_t>(), abort_p);
  HDG_HIP(hipGetLastError());
  static const char* plain_names[] = {"prep_kernel",     "count_kernels",   "coarse_dist+pick", "sub_expand_kernel",
                                      "sub_dist+pick",   "ranges_kernel",   "plan+sync",        "keep_all_kernel",
                                      "adc_scan_kernel", "adc_select_kernel", "final_scan", "final_dist_kernel",
                                      "final_pick_kernel"};
  static const char* mixed_names[] = {"prep_kernel",     "count_kernels",   "coarse_gemm+select+budget",
                                      "sub_expand_kernel", "sub_gemm+select", "ranges_kernel", "plan+sync",
                                      "keep_all_kernel", "adc_scan_kernel", "adc_select_kernel", "final_scan",
                                      "final_dist_kernel", "final_pick_kernel"};
  w.ev_names = mixed ? mixed_names : plain_names;
  w.ev_used = w.capturing ? 0 : evn;
}

}  // namespace hdg
