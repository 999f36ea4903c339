// Super-Grid index: device-resident state, build and query entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "hdg_budget.hpp"
#include "hdg_common.hpp"
#include "hdg_config.hpp"
#include "hdg_grid.hpp"

namespace hdg {

// Label words: a point's label dims packed into one u64, dim j as (label - minv[j]) in the bits
// [off[j], off[j] + popcount(mask[j])), each field followed by a zero guard bit (set in `guard`) so that all
// fields can be compared with the query's bounds by two 64-bit subtractions.  Exact for non-negative integer
// labels (see prepare_label_words).
struct LabelWordLayout {
  uint32_t off[kMaxDims] = {0};
  uint32_t mask[kMaxDims] = {0};
  uint64_t minv[kMaxDims] = {0};
  uint64_t guard = 0;
};

// Device layout.  Every per-point array is in z-pool order: sub-clusters in cluster-major order
// (Super-Grid-IVF_PQ.cu:486-504), points inside a sub-cluster ascending by (Morton code, id)
// (:604-617).  Reading a Morton range therefore reads contiguous memory for codes, labels,
// cells and vectors.
struct IndexState {
  GridShape grid;
  uint64_t N = 0, d = 0, C = 0, CS = 0, M = 0, K = 0, pq_len = 0;
  int l = 0;
  int max_subs_per_cluster = 0;
  uint32_t max_seg_len = 0;
  // host copies kept for save / inspection
  std::vector<uint32_t> h_sub_off;   // [C+1] sub-cluster range of cluster c
  std::vector<uint64_t> h_seg_off;   // [CS+1] z-pool range of sub-cluster s
  std::vector<uint32_t> h_zid;       // [N]
  std::vector<uint64_t> h_zcode;     // [N]
  // device arrays
  DevBuf vec;         // f32 [N][d]
  DevBuf lab;         // f32 [N][l] (original label order)
  DevBuf cells;       // u32 [n_valid][N] cell coordinate of each valid sorted dim
  DevBuf zcode;       // u64 [N]
  DevBuf zid;         // u32 [N] original row id
  DevBuf codes;       // u8  [N][M]
  // the codes again, rows in blocks of 64 (block b at byte b * 64 * M): first the M / 16 sixteen-byte words of
  // the 64 rows (word w of row p at w * 1024 + (p % 64) * 16), then, when M % 16 == 8, their last eight codes
  // (at (M / 16) * 1024 + (p % 64) * 8); so 64 consecutive rows' word is one contiguous piece for the streaming
  // ADC scan.  codes_wb = 16 when M % 8 == 0 (the layout exists), else 0.
  DevBuf codes_blk;
  uint32_t codes_wb = 0;
  DevBuf pq_centers;  // f32 [M][K][pq_len]
  DevBuf centers;     // f32 [C][d]
  DevBuf sub_centers; // f32 [CS][d] cluster-major
  DevBuf sub_off;     // u32 [C+1]
  DevBuf seg_off;     // u32 [CS+1]
  DevBuf cluster_of_sub;  // u32 [CS]
  // mixed-precision coarse scoring (PRECISION 1)
  uint64_t Dp = 0;        // d rounded up to a multiple of 32
  DevBuf centers_bf16;    // bf16 [C][Dp], zero padded
  DevBuf center_norm;     // f32 [C]: |c|^2 of the bf16-rounded center
  DevBuf sub_centers_bf16;  // bf16 [CS][Dp]
  DevBuf sub_center_norm;   // f32 [CS]
  // bucket grid over the valid dims for enumerating a query's in-box points: bucket coordinate
  // b_j = cell_j * nb[j] / Mv[j]; buckets row-major with the last valid dim fastest; entries sorted
  // by (bucket, z-pool position).
  uint32_t Mv[kMaxDims] = {0};      // cells per valid dim
  uint32_t nb[kMaxDims] = {0};      // buckets per valid dim
  uint64_t bstride[kMaxDims] = {0};
  uint64_t Nb = 0;
  DevBuf bstart;      // u32 [Nb+1]
  DevBuf ecell;       // [n_valid][N] cells of the entries: u16 when cells16 (every valid dim <= 65536 cells), else u32
  DevBuf esub;        // u32 [N] sub-cluster (table position) of the entries; packed entries: sub << pk_shift |
                      // (z-pool position - seg_off[sub])
  DevBuf epos;        // u32 [N] z-pool position of the entries (empty when the entries are packed)
  bool cells16 = false;
  bool entries_packed = false;  // bits of CS - 1 plus bits of max_seg_len - 1 fit 32
  uint32_t pk_shift = 0;
  // coarse occupancy sums per sub-cluster for the dense count: coarse cell g_j = cell_j * sg[j] / Mv[j];
  // table [CS][sat_per] of inclusive prefix sums over the (sg + 1)^n_valid grid shifted by one per dim
  // (entry x counts the points with coarse cells < x); sat_per = 0 when the table would be too large
  uint32_t sg[kMaxDims] = {0};
  uint64_t sstride[kMaxDims] = {0};
  uint64_t sat_per = 0;
  DevBuf sat;         // u32 [CS][sat_per]
  // every label value is an integer: then a point passing the exact label predicate lies in the query's
  // cell box, so the in-box points of the chosen sub-clusters are a superset of the keep-all survivors
  bool labels_integral = false;
  // the labels again as one word per point (z-pool order) when they are integers whose columns' ranges fit
  // (lab_words): the label post-filters read 8 bytes per point instead of 4 l
  DevBuf labw;        // u64 [N]
  bool lab_words = false;
  LabelWordLayout lw;
  DevBuf lw_dev;      // the layout again in device memory (the kernels read it once per query)
  bool built = false;
};

// A saved index state read into host memory and checked (hdg_io.cpp): the host part of IndexState plus the
// arrays that go to the device.
struct HostCheckpoint {
  IndexState S;
  std::vector<float> centers, subc, pqc, vec, lab;
  std::vector<uint8_t> codes;
  std::vector<uint32_t> cells;
};
// Reads and validates the checkpoint in `dir` for an index with l_expected label dims; throws kIoError on
// any inconsistency (shapes, offsets, permutations, value ranges).  Touches no device.
HostCheckpoint read_checkpoint(const std::string& dir, int l_expected);

// Bucket grid on the device from the z-pool-ordered cells (st.cells, [n_valid][N]) and sub-cluster offsets
// (st.seg_off); sets the bucket fields.
void build_bucket_grid(IndexState& st, hipStream_t s);

// Balanced hierarchical k-means (RAFT kmeans_balanced::fit_predict, Super-Grid-IVF_PQ.cu:466-476) of n rows
// X [n][w] (device) into k clusters: centers [k][w] and labels [n] returned on the host.
void kmeans_balanced(const float* X, uint64_t n, uint32_t w, uint32_t k, uint32_t n_iters,
                     std::vector<float>& centers, std::vector<uint32_t>& labels, hipStream_t s);
// Fine cluster counts of the mesoclusters (sizes meso_sizes, n rows in all) for k clusters.
std::vector<uint32_t> arrange_fine_clusters(uint32_t k, const std::vector<uint64_t>& meso_sizes, uint64_t n);

// Sets labels_integral from h_lab (the labels in any row order, host) and, when every label is a non-negative
// integer of at most 2^53, every column's range max - min is below 2^31 and the columns' bit widths (at least
// one each) plus one guard bit per column sum to at most 64, packs st.lab (z-pool order, device) into st.labw
// and sets lab_words.
void prepare_label_words(IndexState& st, const std::vector<float>& h_lab, hipStream_t s);

// Fills codes_blk / codes_wb from the row-layout codes (device).
void prepare_blocked_codes(IndexState& st, hipStream_t s);

// Fills the bf16 center copies, their norms and Dp from host centers [C][d] and sub-centers [CS][d].
void prepare_mixed_centers(IndexState& st, const std::vector<float>& centers, const std::vector<float>& sub_centers,
                           hipStream_t s);

class Index {
 public:
  Index(const Config& cfg, const FilterConfig& fc, int device);
  ~Index();

  // Super-Grid-IVF_PQ.cu:420-646 (index::build).  dataset [n][d], labels [n][label_cols] are
  // device pointers; label_cols must equal filter.conf's l.
  void build(const float* dataset, uint64_t n, uint64_t d, const float* labels, uint64_t label_cols,
             hipStream_t stream);
  // Super-Grid-IVF_PQ.cu:985-1248 (index::query) on nq queries; reference batches of
  // `parity_batch` queries share beta.  Outputs are device pointers [nq][topk].
  void query(const float* queries, uint64_t nq, uint64_t d, const float* qlabels, uint64_t qcols,
             float* out_dist, uint64_t* out_ids, hipStream_t stream);

  // query() with its whole workspace carved from this device's fixed scratch array (kNoAllocScratchBytes):
  // no device or pinned allocation, no graph capture, no stage timing.  Chunks of whole reference batches are
  // halved until one fits; a batch that does not fit, or call-wide counters beyond a quarter of the array, fail.
  // Calls that overlap on one device race on the shared array.
  void query_noalloc(const float* queries, uint64_t nq, uint64_t d, const float* qlabels, uint64_t qcols,
                     float* out_dist, uint64_t* out_ids, hipStream_t stream);

  void save(const std::string& dir) const;
  void load(const std::string& dir, hipStream_t stream);

  const Config& config() const { return cfg_; }
  Config& config() { return cfg_; }
  const FilterConfig& filter() const { return fc_; }
  const IndexState& state() const { return st_; }
  int device() const { return device_; }

  // Per-query counters of the last query call, columns (C_q Morton-range candidates, in-box count of
  // the chosen sub-clusters, vectors re-ranked exactly, centers + sub-centers scored, candidates scanned
  // with PQ ADC, PQ keep width); out[c] may be null.  Synchronises with the device.
  void last_stats(uint64_t nq, uint64_t* const out[6]) const;
  // Chosen sub-cluster ids of the last query call, [nq][sub_lists], UINT32_MAX past each query's end.
  void last_chosen_subs(uint64_t nq, uint32_t* out) const;
  // Per-stage device time (ms, HIP events) of the last query chunk, for profiling.
  std::vector<std::pair<std::string, float>> stage_times() const;

  // learned per-query budget (hdg_budget.hip)
  bool budget_model_ready() const;
  bool budget_capturing() const;                 // budget_fit is recording feature rows
  bool budget_probes_all(uint32_t C) const;      // every level of the loaded model probes >= C clusters
  // Per-query sub-cluster budgets from the MLP; trims n_chosen_s (chosen sub-clusters, sorted) to them.
  void run_budget_model(uint64_t qn, int l, const float* ranges, const uint32_t* box_lo, const uint32_t* box_w,
                        int n_valid, const float* feat_dist, uint32_t* n_chosen_s, uint32_t* budget_T,
                        uint32_t* budget_S, hipStream_t s, const uint32_t* abort_flag = nullptr);
  void capture_features(uint64_t q0, uint64_t qn, int l, const float* ranges, const uint32_t* box_w, int n_valid,
                        const float* feat_dist, hipStream_t s);
  // Labels every training query with its minimal budget level (traces of this pipeline against the
  // caller's ground truth ids, host [n][k]), then trains the MLP on the GPU.
  void budget_fit(const float* queries, uint64_t n, uint64_t d, const float* qlabels, uint64_t qcols,
                  const uint64_t* gt_host, uint64_t k, const BudgetFitParams& p, hipStream_t s);
  const BudgetReport& budget_report() const;
  // The raw (unstandardised) feature rows [n][48] the PRECISION 1 pipeline computes for these queries.
  void budget_debug_features(const float* queries, uint64_t n, uint64_t d, const float* qlabels, uint64_t qcols,
                             float* out_host, hipStream_t s);
  void budget_save(const std::string& path) const;
  void budget_load(const std::string& path, hipStream_t s);
  void budget_free();
  void debug_coarse_scores(const float* queries, uint64_t nq, float* out_host, hipStream_t s);
  // Coordinate descent over the launch knobs (ADC_PER, ADC_WGS, QUERY_CHUNK, SCORE_BLOCK, ADC_SPLIT, GRAPH;
  // SPARSE_MAX_TESTS under PRECISION 0 only) timed on these queries (hdg_autotune.cpp); keeps the fastest setting whose
  // results equal the starting one's bit for bit and returns the trial log.
  std::string autotune(const float* queries, uint64_t nq, uint64_t d, const float* qlabels, uint64_t qcols,
                       float* out_dist, uint64_t* out_ids, uint32_t reps, hipStream_t s);
  struct BudgetModel;

 private:
  BudgetModel* budget_ = nullptr;
  uint64_t graph_epoch_ = 0;  // bumped by build / load: captured graphs of an older index state are stale
  uint64_t config_digest() const;
  void drop_graphs();
  void query_chunk(const float* queries, uint64_t q0, uint64_t nq, const float* qlabels,
                   uint64_t qcols, uint64_t group, float* out_dist, uint64_t* out_ids,
                   hipStream_t stream);
  // argument checks shared by query() and query_noalloc(); returns the reference batch size G
  uint64_t check_query(const float* queries, uint64_t nq, uint64_t d, const float* qlabels, uint64_t qcols,
                       float* out_dist, uint64_t* out_ids) const;
  Config cfg_;
  FilterConfig fc_;
  int device_;
  IndexState st_;
  // device copies of the filter configuration
  DevBuf f_type_, f_shift_, f_map_, f_div_;
  uint64_t f_map_len_ = 0;
  // query workspace (grown on demand, reused across calls)
  struct Work;
  Work* w_;
  Work* wn_ = nullptr;      // workspace of query_noalloc (every buffer a piece of the scratch array)
  Work* last_w_ = nullptr;  // workspace of the last query call (read by last_stats / last_chosen_subs)
  void* scratch_ = nullptr; // device address of this device's scratch array
};

// Size of the per-device scratch array behind Index::query_noalloc.
constexpr uint64_t kNoAllocScratchBytes = 64ull << 20;

// findMaxFactor (math_utils.h:40-60).
int find_max_factor(uint64_t a, uint64_t b);

}  // namespace hdg
