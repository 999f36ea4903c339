/*
 * hdgrid.h -- C ABI of the MI355X-native Super-Grid (HDGrid) filtered-ANN search library.
 *
 * The entry points replace the reference's C++ operator API for the Super-Grid path
 * (reference tree: wangzheng456123/HDGrid, paths relative to its root):
 *   index_base_t::build / index_base_t::query      src/core/interfaces.cuh:27-46
 *   super_grid::index::build                       src/index/2-level_index/Super-Grid-IVF_PQ.cu:420-646
 *   super_grid::index::query                       src/index/2-level_index/Super-Grid-IVF_PQ.cu:985-1248
 *   index_factory::create / config_factory::create src/index/index-factory.cuh:12-38,40-89
 *   HDGrid.conf and filter.conf parsing            src/utils/config_utils.h:12-194
 *
 * Conventions.
 *   - Array arguments of hdg_build / hdg_query are DEVICE pointers (HIP), row-major, on the index's
 *     device, exactly like the reference's raft::device_matrix_view inputs and caller-allocated
 *     outputs (interfaces.cuh:6-25).
 *   - `stream` is a hipStream_t passed as void* (NULL = the default stream).
 *   - Every call returns an hdg_status; on failure hdg_last_error() describes the error (thread
 *     local).  The library never calls exit().
 *   - Result ids are row indices of the dataset given to hdg_build; results are sorted by ascending
 *     distance.  Missing results: id 2^32 (l == 1) or UINT64_MAX with distance 1e6 (EXPS0 > 1), or
 *     id UINT64_MAX with distance FLT_MAX (EXPS0 <= 1) -- the reference's pad values.
 */
#ifndef HDGRID_H
#define HDGRID_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hdg_index hdg_index;
typedef int hdg_status;

#define HDG_OK 0
#define HDG_EINVAL 1       /* invalid argument or input the reference cannot represent */
#define HDG_ECONFIG 2      /* HDGrid.conf / filter.conf error, unknown algorithm */
#define HDG_EIO 3          /* file I/O */
#define HDG_EHIP 4         /* HIP runtime error */
#define HDG_ENOMEM 5       /* device allocation failed */
#define HDG_ESTATE 6       /* call order (query before build, ...) */
#define HDG_EUNSUPPORTED 7 /* outside the supported envelope (d > 8192, clusters > 4096, ...) */

/* Thread-local description of the last failure of a call on this thread. */
const char* hdg_last_error(void);

/* Creates an index for `algorithm` ("super-grid", case-insensitive) from an HDGrid.conf and a
 * filter.conf on HIP device `device`.  filter_conf_path may be NULL: then <PATH>filter.conf is used,
 * as the engine does (search-engine-auto-batch.cu:134-141).
 * Replaces index_factory::create + config_factory::create (index-factory.cuh:12-38,40-89) and the
 * super_grid::index constructor (Super-Grid-IVF_PQ.cuh:84-87). */
hdg_status hdg_create(const char* algorithm, const char* conf_path, const char* filter_conf_path,
                      int device, hdg_index** out);

/* Same as hdg_create with the two configuration files given as text.  No reference counterpart:
 * for hosts that hold configuration in memory. */
hdg_status hdg_create_from_text(const char* algorithm, const char* conf_text, const char* filter_text,
                                int device, hdg_index** out);

/* Parses an HDGrid.conf text and a filter.conf text without touching any device and writes the parsed
 * values as one JSON object into out (cap bytes).  No reference counterpart: configuration checking
 * for hosts and tools (the reference parses inside its constructors, config_utils.h:33-172). */
hdg_status hdg_validate_config_text(const char* conf_text, const char* filter_text, char* out, uint64_t cap);

/* Builds (or rebuilds) the index from dataset [n][dim] f32 and labels [n][label_cols] f32 (device
 * pointers; label_cols must equal filter.conf's l).  The index copies what it needs; the inputs may
 * be freed afterwards.
 * Replaces super_grid::index::build(build_input_t const&) (Super-Grid-IVF_PQ.cu:420-646). */
hdg_status hdg_build(hdg_index* index, const float* dataset, uint64_t n, uint64_t dim,
                     const float* labels, uint64_t label_cols, void* stream);

/* Filtered top-k search of queries [nq][dim] f32 with query_labels [nq][qlabel_cols] f32
 * (qlabel_cols = l * FILTER_DIM).  Writes out_dist [nq][topk] f32 and out_ids [nq][topk] u64 (device
 * pointers).  Queries are processed in reference batches of PARITY_BATCH (default
 * findMaxFactor(125, nq), search-engine-auto-batch.cu:484-486) that share beta.
 * Replaces super_grid::index::query(query_input_t const&, query_output_t&)
 * (Super-Grid-IVF_PQ.cu:985-1248). */
hdg_status hdg_query(hdg_index* index, const float* queries, uint64_t nq, uint64_t dim,
                     const float* query_labels, uint64_t qlabel_cols, float* out_dist,
                     uint64_t* out_ids, void* stream);

/* hdg_query with no allocation: no counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * results and errors as hdg_query, but the whole workspace is carved from one fixed 64 MiB array per device
 * (a static device array the library holds until the process ends; its address is looked up once per device
 * when an index is created): the call makes no hipMalloc / hipFree / pinned allocation, never captures a
 * graph (GRAPH is ignored) and records no stage times.  Queries run in chunks of whole PARITY_BATCH groups,
 * halved until one chunk's workspace fits.  Limits: the call-wide counters (nq * (48 + 4 * SUB_LISTS) bytes)
 * must fit in a quarter of the array, else HDG_EUNSUPPORTED; one PARITY_BATCH group whose workspace exceeds the
 * array fails with HDG_ENOMEM.  Every call on a device shares that device's array and overwrites it: calls
 * that overlap on one device (other threads, other streams, other indexes) race, and hdg_last_query_stats /
 * hdg_last_chosen_subs after this call read the array, so they are valid until the next allocation-free call
 * on the device. */
hdg_status hdg_query_noalloc(hdg_index* index, const float* queries, uint64_t nq, uint64_t dim,
                             const float* query_labels, uint64_t qlabel_cols, float* out_dist,
                             uint64_t* out_ids, void* stream);

/* Releases the index and its device memory (the reference frees MEM_INDEX, mmr.cuh:73-80). */
hdg_status hdg_destroy(hdg_index* index);

/* Applies one HDGrid.conf `key = value` (same grammar as the file, config_utils.h:33-88). */
hdg_status hdg_set_param(hdg_index* index, const char* key, const char* value);
/* Reads an integer HDGrid.conf key. */
hdg_status hdg_get_param(const hdg_index* index, const char* key, uint64_t* value);

/* Writes the index state to directory `dir` (.npy arrays + meta.json) / reads it back.  No reference
 * counterpart (the reference rebuilds the index per data batch): checkpoint/resume and the shared
 * state that the CPU oracle is run on. */
hdg_status hdg_save(const hdg_index* index, const char* dir);
/* hdg_load checks the whole state before it replaces the index: every array's shape against meta.json,
 * monotone offsets spanning the rows, zid a permutation, Morton codes ascending per sub-cluster, cells inside
 * their dims, PQ codes below the codeword count; HDG_EIO otherwise (the index is left as it was). */
hdg_status hdg_load(hdg_index* index, const char* dir, void* stream);
/* The same checks without an index or a device: HDG_OK when `dir` holds a consistent state for
 * `label_dims` label dims, HDG_EIO otherwise. */
hdg_status hdg_validate_checkpoint(const char* dir, int label_dims);

/* Per-query counters of the last hdg_query (host arrays of nq entries; any may be NULL):
 * cand = Morton-range candidates C_q (Super-Grid-IVF_PQ.cu:1184), real = in-box count of the chosen
 * sub-clusters (:1199-1210), refined = vectors re-ranked exactly, scored = centers + sub-centers
 * scored exactly, adc = candidates scanned with PQ, kprime = PQ keep width. */
hdg_status hdg_last_query_stats(const hdg_index* index, uint64_t nq, uint64_t* cand, uint64_t* real,
                                uint64_t* refined, uint64_t* scored, uint64_t* adc, uint64_t* kprime);

/* Chosen sub-clusters of the last hdg_query's last chunk, [nq][SUB_LISTS] u32 host array in
 * z-pool sub-cluster numbering, UINT32_MAX padded (Super-Grid-IVF_PQ.cu:1046-1056 output). */
hdg_status hdg_last_chosen_subs(const hdg_index* index, uint64_t nq, uint32_t* out);

/* Stage timings (ms) of the last hdg_query chunk: names joined by ',' into `names` (cap bytes),
 * values into ms[0..*n). */
hdg_status hdg_last_stage_ms(const hdg_index* index, char* names, uint64_t cap, float* ms, uint64_t* n);

/* Learned per-query probe budget (used when PRECISION = 1 and BUDGET_MODEL = 1).  No reference
 * counterpart: the reference probes a fixed n_list / sub_lists for every query.
 * hdg_budget_fit labels each of the n training queries (device arrays, like hdg_query) with the
 * smallest budget level whose top-k reaches `target` recall against gt_ids (host [n][TOPK] u64, invalid
 * ids >= 2^63), running this pipeline at every level, then trains the 48-64-64-1 MLP on the device.
 * Ladder: level i keeps round(s0*ratio^i) sub-clusters (capped by SUB_LISTS) of all non-zero clusters; t0
 * must be 0 (cluster budgets are not modelled: HDG_EUNSUPPORTED otherwise).  Labels come from the plain path
 * (BUDGET_LABELS = 0, the default) or the mixed one (1).  Features per query: filter widths, grid-cell extents
 * and the 32 smallest BF16 coarse-centroid scores among its non-zero clusters (+ |q|^2); at query time the
 * budget trims the score-sorted sub-cluster choice.  The margin added to the prediction is the smallest one at
 * which the deployed path (PRECISION 1 with the model) reaches `target` mean recall plus one standard error of
 * that mean on the held-out 10% of the training queries.  report receives 8 floats: final loss, margin, holdout
 * recall, holdout mean level, mean label, recall of the top level, n_train, n_holdout. */
hdg_status hdg_budget_fit(hdg_index* index, const float* queries, uint64_t n, uint64_t dim,
                          const float* query_labels, uint64_t qlabel_cols, const uint64_t* gt_ids, uint64_t k,
                          int levels, double t0, double s0, double ratio, float target, int epochs, float lr,
                          uint64_t seed, float* report, void* stream);
hdg_status hdg_budget_save(const hdg_index* index, const char* path);
hdg_status hdg_budget_load(hdg_index* index, const char* path, void* stream);
/* One forward/backward pass of the training kernel on n samples (X [n][48], Y [n], params [P], all host);
 * writes the parameter gradients (grads [P]) and the batch loss.  For checking the HIP backward against a
 * CPU implementation.  P = hdg_budget_param_count(). */
hdg_status hdg_budget_debug_grads(const float* params, const float* X, const float* Y, uint32_t n, float* grads,
                                  float* loss);
int hdg_budget_param_count(void);
/* The budget model's raw (unstandardised) input rows for n queries (device arrays, as hdg_query), written to
 * out [n][48] (host): log1p of each label dim's filter width (8), log1p(extent + 1) of each valid grid dim's cell
 * box (8), the 32 smallest BF16 coarse scores |c|^2 - 2<q,c> among the query's non-zero clusters plus |q|^2 (32).
 * Runs the PRECISION 1 pipeline once.  For checking the feature and inference kernels. */
hdg_status hdg_budget_debug_features(hdg_index* index, const float* queries, uint64_t n, uint64_t dim,
                                     const float* query_labels, uint64_t qlabel_cols, float* out, void* stream);

/* Coarse scores of the PRECISION 1 path for nq queries (device array): out [nq][CLUSTERS] host f32 with
 * out[q][c] = |c|^2 - 2 <q, c> computed by the BF16 MFMA kernel.  For checking that kernel. */
hdg_status hdg_debug_coarse_scores(hdg_index* index, const float* queries, uint64_t nq, float* out, void* stream);

/* Budget-model coarse-score features of nq queries (device arrays): feat [nq][32] = the 32 smallest of
 * scores[q][list[q][0 .. n_list[q])] ascending (the last repeated when fewer; 0 when the list is empty), each
 * plus |queries[q]|^2 (queries [nq][dim]; list rows cap apart, cap <= 512; scores rows ld apart) -- the features
 * the query path computes when every non-zero cluster is probed (the learned budget replaces the fixed
 * SUB_LISTS choice of Super-Grid-IVF_PQ.cu's search; its features have no counterpart in the reference).
 * For checking that kernel on hand-made lists and ties. */
hdg_status hdg_debug_coarse_features(const float* queries, uint64_t nq, uint64_t dim, const uint32_t* list,
                                     uint32_t cap, const uint32_t* n_list, const float* scores, uint64_t ld,
                                     float* feat, int device, void* stream);

/* Per-query top-k of scored lists (device arrays), as the query path picks clusters and sub-clusters from
 * their BF16 scores (replaces select_sub_clusters' sort of Super-Grid-IVF_PQ.cu:885-896 on those scores):
 * query q's items are list[q][0 .. n_list[q]) (rows cap apart) scored scores[q][item] (rows ld apart); k_q =
 * min(k, budget[q]) when budget (nullable) gives a non-zero value; chosen [nq][out_ld] gets the list positions
 * of the min(n, k_q) smallest scores, ascending, ties by position, n_chosen [nq] their count; feat (nullable,
 * [nq][32]) the 32 smallest + |queries[q]|^2.  wave = 1 runs the wave-per-query kernel (lists <= 512), 0 the
 * workgroup kernel.  Synchronous.  For checking the two selection kernels against each other and a sort. */
hdg_status hdg_debug_scored_select(const float* queries, uint64_t nq, uint64_t dim, const uint32_t* list,
                                   uint32_t cap, const uint32_t* n_list, const float* scores, uint64_t ld,
                                   uint32_t k, const uint32_t* budget, uint32_t out_ld, uint32_t* chosen,
                                   uint32_t* n_chosen, float* feat, int wave, int device, void* stream);

/* Balanced hierarchical k-means of n rows (device array rows [n][w]) into k clusters, as the index build
 * groups its clusters*sub_clusters fine centers into `clusters` groups (replaces the RAFT
 * kmeans_balanced::fit_predict call at Super-Grid-IVF_PQ.cu:466-476; n_iters = BALANCE_ITERS there).
 * Writes centers [k][w] and labels [n] (host).  For checking the build's grouping. */
hdg_status hdg_debug_kmeans_balanced(const float* rows, uint64_t n, uint32_t w, uint32_t k, uint32_t n_iters,
                                     float* centers, uint32_t* labels, int device, void* stream);

/* Top-k merge of per-data-batch results (device arrays): dist/ids [n_batches][nq][topk], ids local to
 * their batch; batch b holds rows [b*data_batch_size, min(n_data, (b+1)*data_batch_size)).  Writes
 * out_dist/out_ids [nq][topk] (device): per query the topk smallest distances, ascending, ties by (batch,
 * rank); ids become start_offset + b*data_batch_size + local, pads (ids >= the batch's row count) are kept.
 * topk <= 4096.  Replaces merge_intermediate_result (merge_res.cu:18-55), which adds the offsets as
 * floats (to pads too) and reads the batches back from files. */
hdg_status hdg_merge_topk(const float* dist, const uint64_t* ids, uint64_t n_batches, uint64_t data_batch_size,
                          uint64_t n_data, uint64_t start_offset, uint64_t nq, uint64_t topk, float* out_dist,
                          uint64_t* out_ids, void* stream);

/* Memory-predictor fit: samples[i] = {data rows, query rows, device bytes used} of four probe runs; solves
 * c0*d + c1*q + c2*d*q + c3 = used by the reference's gauss (math_utils.h:6-41; calc_mem_coeff,
 * search-engine-auto-batch.cu:381-428) and writes coeff[4].  Returns gauss's code: 0 unique solution,
 * 1 rank deficient, 2 inconsistent.  Host only. */
int hdg_fit_mem_coeff(const double samples[4][3], double coeff[4]);

/* Launch-parameter autotuner.  Times full queries of the given batch (device arrays, as hdg_query;
 * out_dist/out_ids receive the results) under a ladder of values of each result-invariant knob
 * (ADC_PER, ADC_WGS, QUERY_CHUNK, SCORE_BLOCK, ADC_SPLIT, GRAPH, and SPARSE_MAX_TESTS under PRECISION 0 only:
 * under PRECISION 1 it changes which queries refine from their in-box lists), one knob at a time,
 * `reps` timed runs per value (median), and keeps the fastest value whose ids and distances are bit-identical to those of the
 * starting configuration.  The chosen values stay set on the index; the trial log is written to
 * report (cap bytes, may be NULL).  No reference counterpart: the reference's launch shapes are
 * compile-time constants. */
hdg_status hdg_autotune(hdg_index* index, const float* queries, uint64_t nq, uint64_t dim,
                        const float* query_labels, uint64_t qlabel_cols, float* out_dist, uint64_t* out_ids,
                        uint32_t reps, char* report, uint64_t cap, void* stream);

/* Memory-bound batch predictor: split_task (search-engine-auto-batch.cu:430-489) with the
 * reference's model c0*d + c1*q + c2*d*q + c3 <= min(available/5, mem_bound). */
hdg_status hdg_split_task(const double coeff[4], uint64_t n_data, uint64_t n_queries, uint64_t topk,
                          uint64_t mem_bound, uint64_t available, uint64_t data_width,
                          uint64_t* query_batch, uint64_t* data_batch);

/* findMaxFactor (math_utils.h:40-60): largest factor of b below a (b itself when b <= a). */
int hdg_find_max_factor(uint64_t a, uint64_t b);

#ifdef __cplusplus
}
#endif

#endif /* HDGRID_H */
