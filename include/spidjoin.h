/*
 * spidjoin.h -- C ABI of libspidjoin.so (MI355X / gfx950).
 *
 * The drop-in boundary of the reference is the C++ class
 * pidjoin::JoinInstance (include/join_internals.hpp:301-323 of the reference,
 * implemented in src/query.cc).  The reference has no C or FFI layer; this
 * header is the plain-C form of that same surface (plain pointers and sizes,
 * no C++ or torch types), on which this library's own C++ JoinInstance
 * (pidjoin.hpp) is built and which ctypes / cgo / JNI stubs can bind.
 *
 * Conventions
 *   - Tuples are 8-byte {uint32 key; uint32 rid} (kv_pair_t,
 *     type_defs.hpp:53-56 of the reference).  Key 0 means "empty" and never
 *     joins (glb_partition_count.c:192).  R keys must be unique (PK side).
 *   - Output pairs are {uint32 r_rid; uint32 s_rid}, one per matching S tuple
 *     (a2a_srj_probe_inner_mt.c:142-143), in no particular order, with an
 *     exact count (no zero padding).
 *   - Every function returns 0 on success or a negative SPJ_E* code; the
 *     message of the last error on this thread is spj_last_error().  Nothing
 *     calls exit().  (The reference printf()s and exit(-1)s instead.)
 */
#ifndef SPIDJOIN_H
#define SPIDJOIN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { uint32_t key; uint32_t rid; } spj_tuple_t;
typedef struct { uint32_t r_rid; uint32_t s_rid; } spj_pair_t;

enum {
    SPJ_OK = 0,
    SPJ_EINVAL = -1,       /* bad argument */
    SPJ_EHIP = -2,         /* HIP runtime error */
    SPJ_ENOMEM = -3,       /* device or host allocation failed */
    SPJ_EDUPKEY = -4,      /* duplicate R key (a2a_srj_build.c:265-271) */
    SPJ_EOVERFLOW = -5,    /* output capacity exceeded */
    SPJ_ESTATE = -6,       /* call out of order (e.g. FinishJoin before StartJoin) */
    SPJ_EBUSY = -7         /* admission refused (serving HBM budget) */
};

/* Join algorithm (JOIN_DESIGN_JOIN_ALGORITHM_*, idpHandler.hpp:126-129). */
enum { SPJ_ALG_AUTO = -1, SPJ_ALG_PHJ = 0, SPJ_ALG_NPHJ = 1, SPJ_ALG_SMJ = 2, SPJ_ALG_NL = 3 };

/* Tunables; -1 = automatic (tuner table, else built-in defaults). */
typedef struct {
    int32_t algorithm;        /* SPJ_ALG_*: AUTO = NPHJ when no partition parameter below is
                                 set, no tuner entry covers the instance, and the call is
                                 small: one join with |R| <= 2^21 and |R| |S| <= 2^44, or a
                                 batch whose totals have |R| <= 2^20 and |R| |S| <= 2^44;
                                 else hash, nested loop for tiny partitions;
                                 NPHJ = no partition pass: one open-addressing table of
                                 >= 2|R| 8-byte slots in device memory, probed by every S tuple */
    int32_t radix_bits;       /* partition-pass radix bits (0..12) */
    int32_t table_log2;       /* LDS hash-table buckets, log2 (6..13; 12 = the reference's 4096) */
    int32_t heavy;            /* heavy-key skew handling: 0 off, 1 on; AUTO = on for |S| >= 2^20 (in a
                                 batch: >= 2^14) */
    int32_t heavy_min_count;  /* summed sample count that makes a key heavy */
    int32_t scatter_ipt;      /* scatter tuples per thread per round (1024 threads): 4, 8, 10 or 12 */
    int64_t tile_tuples;      /* tuples per partition-pass tile of a counted relation (cells: two
                                 scatter rounds, 2 * 1024 * scatter_ipt) */
    int64_t items_s_max;      /* S tuples per join work item, at most: a call or batch whose S is
                                 small uses smaller items, about two per CU (no fewer than 4096
                                 S tuples each), so that its partitions are joined side by side */
    int32_t part_layout;      /* partitions of R and S: 0 counted (a histogram pass sizes them
                                 exactly), 1 cells (no histogram pass: fixed-capacity cells per
                                 scatter group and digit; the join is redone counted when a cell
                                 overflows); AUTO = cells for R of a hash or sort-merge join, and for its S
                                 from 2^14 tuples with heavy-key handling on */
    int32_t cell_capacity_pct;/* cells: capacity in % of the mean fill (1..200); AUTO = 150% + 8 sigma + 64 */
    int32_t radix_passes;     /* radix partition passes, 1 or 2 (single-instance calls only).  Two:
                                 the first pass partitions R and S by radix_bits of
                                 SuperFastHash (in the layout part_layout picks; heavy keys joined in
                                 its S pass; scatter_ipt applies to it), the second (with its own
                                 round size) splits every partition by the next bits
                                 (chosen per partition so R partitions fit one LDS table) and
                                 joins it.  AUTO = 2 when one pass of 12 bits would leave R
                                 partitions of 4 or more LDS tables (|R| > 3 * 2^26), else 1 */
} spj_config_t;

/* Per-phase device time of the last call (HIP events on its stream) and, for
 * single-instance calls, how many tuples went through the partitions (heavy
 * S tuples are joined in the scatter and never partitioned). */
typedef struct {
    float heavy_ms, hist_ms, scan_ms, scatter_ms, items_ms, join_ms, total_ms;
    uint32_t heavy_keys;
    uint64_t r_partitioned, s_partitioned;
    uint32_t s_cells;   /* 1: S went to fixed-capacity cells (no histogram pass), 0: counted */
    uint32_t r_cells;   /* the same for R */
    uint32_t passes;    /* radix partition passes the call ran (hist, scan and scatter sum both) */
} spj_phase_times_t;

void spj_config_default(spj_config_t *cfg);
const char *spj_last_error(void);

/* Device selection and facts (hipGetDeviceProperties). */
int spj_init(int device);
int spj_device_info(int *cus, uint64_t *lds_per_block, char *arch, int arch_len);

/* Device memory helpers, so hosts without a GPU framework can drive the
 * library (the reference's SEND / RECV host<->device moves,
 * queryHostFunc.cc:1058-1126, 1421-1498). */
int spj_dev_alloc(void **ptr, uint64_t bytes);
int spj_dev_free(void *ptr);
int spj_h2d(void *dst, const void *src, uint64_t bytes);
int spj_d2h(void *dst, const void *src, uint64_t bytes);
int spj_dev_memset(void *ptr, int value, uint64_t bytes);
/* Device-to-device copy; returns the copy's device time in *ms when non-NULL
 * (used to calibrate achievable HBM bandwidth). */
int spj_d2d(void *dst, const void *src, uint64_t bytes, float *ms);
int spj_dev_sync(void);
/* Launch sequences of the device-buffer calls (spj_join_device,
 * spj_join_batch_device) replayed as HIP graphs: a call that repeats the
 * previous call exactly (same pointers, sizes and parameters) replays the
 * previous call's kernels and copies as one captured graph instead of
 * launching them one by one; any other call launches directly.  Timed calls
 * (times != NULL) always launch directly.  The environment variable
 * SPJ_GRAPHS=0 turns replays off.  No counterpart in the reference; these
 * counters say how often a graph was captured and replayed. */
int spj_graph_stats(uint64_t *captures, uint64_t *replays);
/* Order-independent checksum of device-resident pairs: sum and xor of a
 * 64-bit mix of (r_rid << 32 | s_rid) -- compares result multisets in place. */
int spj_pairs_checksum_device(const spj_pair_t *pairs, uint64_t n, uint64_t *sum, uint64_t *xr);

/* One join on device-resident relations: the whole plan of
 * execution.hpp:15-157 (partition R, partition S, join) on one stream.
 * out must hold out_cap >= nS pairs.  times may be NULL. */
int spj_join_device(const spj_tuple_t *R, uint64_t nR, const spj_tuple_t *S, uint64_t nS, spj_pair_t *out,
                    uint64_t out_cap, uint64_t *out_count, const spj_config_t *cfg, spj_phase_times_t *times);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): spj_join_device's no-partition join (SPJ_ALG_NPHJ) with
 * its table, descriptors and control words in a static device array of
 * 16 MiB + 512 bytes that each device holds from library load until the
 * process ends.  Nothing is allocated.  |R| <= 2^20 (SPJ_NOALLOC_MAX_R):
 * larger R is refused with SPJ_EINVAL.  Every call on a device shares that
 * device's array; calls in one process take turns (a lock), and the array
 * is set up again by each call.  out must hold out_cap >= nS pairs; runs on
 * the null stream and returns when the join is done. */
#define SPJ_NOALLOC_MAX_R (1u << 20)
int spj_join_device_noalloc(const spj_tuple_t *R, uint64_t nR, const spj_tuple_t *S, uint64_t nS, spj_pair_t *out,
                            uint64_t out_cap, uint64_t *out_count);

/* The global partition pass alone, on one device-resident relation
 * (COMPOUND_FUNC_GLB_PARTITION, queryCompoundFunc.cc:814-872: GLB_PARTITION_COUNT,
 * the page histogram, GLB_PARTITION_PACKET).  Tuple t goes to partition
 * glb_partition_hash(t.key) & (2^radix_bits - 1) (glb_partition_count.c:192-195);
 * key-0 tuples are skipped.  out (device, n tuples) receives the partitions
 * back to back, partition p at [pstart[p], pstart[p + 1]) in no particular
 * order inside; pstart (device) receives 2^radix_bits + 1 offsets, the last
 * one the number of non-zero keys.  radix_bits 0..12. */
int spj_partition_device(const spj_tuple_t *rel, uint64_t n, int radix_bits, spj_tuple_t *out, uint32_t *pstart);

/* Same on host buffers (H2D, join, D2H of exactly out_count pairs). */
int spj_join_host(const spj_tuple_t *R, uint64_t nR, const spj_tuple_t *S, uint64_t nS, spj_pair_t *out,
                  uint64_t out_cap, uint64_t *out_count, const spj_config_t *cfg);

/* Many independent joins in one launch sequence and one persistent join
 * kernel fed by a device work queue.  Arrays have n entries; out[i] holds
 * nS[i] pairs; out_count[i] receives instance i's exact count.  cfg applies
 * to all instances (per-instance parameters come from the tuner when the
 * fields are -1).  No counterpart in the reference (one join per
 * JoinInstance). */
int spj_join_batch_device(int n, const spj_tuple_t *const *R, const uint64_t *nR, const spj_tuple_t *const *S,
                          const uint64_t *nS, spj_pair_t *const *out, uint64_t *out_count, const spj_config_t *cfg,
                          spj_phase_times_t *times);

/* ---- reference-shaped instance API (pidjoin::JoinInstance, query.cc) ---- */
typedef struct spj_instance spj_instance_t;

/* join_design_t (idpHandler.hpp:131-145); fields without a GPU meaning
 * (rank_num, zipf_factor, rs_ratio, input_size) are accepted and ignored. */
typedef struct {
    int64_t input_size;
    int32_t rank_num;
    int32_t zipf_factor;
    int32_t rs_ratio;
    int32_t rank_set;   /* > 1: rank-set replication -> heavy-key handling on */
    int32_t bank_set;   /* > 1: bank-set replication -> heavy-key handling on */
    int32_t join_alg;   /* 0 phj, 1 nphj (one HBM table, no partition pass), 2 smj, 3 nl */
} spj_join_design_t;

void spj_join_design_default(spj_join_design_t *d);   /* idpHandler.hpp:133-144 defaults */

/* JoinInstance(num_rank_allocated, num_rankgroup, num_rank_in_rankgroup), query.cc:842 */
spj_instance_t *spj_create(int num_rank_allocated, int num_rankgroup, int num_rank_in_rankgroup);
/* LoadColumn(data, n, "left" | "right"), query.cc:213: the caller keeps data alive until spj_finish. */
int spj_load_column(spj_instance_t *inst, const void *data, int64_t num_tuples, const char *col_name);
/* PrepareJoin, query.cc:178: allocate device buffers. */
int spj_prepare(spj_instance_t *inst, const spj_join_design_t *design);
/* StartJoin, query.cc:188: asynchronous -- enqueue H2D + kernels and return. */
int spj_start(spj_instance_t *inst, const spj_join_design_t *design);
/* FinishJoin, query.cc:193: wait, copy exactly *count pairs into a library-
 * allocated host buffer the caller frees with spj_free_result. */
int spj_finish(spj_instance_t *inst, spj_pair_t **pairs, uint64_t *count);
/* FinishJoin's buffers (ResultBuffers_t, type_defs.hpp:82), by layout:
 *   SPJ_RESULT_EXACT (default): one buffer of exactly `count` pairs;
 *   SPJ_RESULT_RANK_PADDED: the reference's gathered layout
 *     (queryHostFunc.cc:1058-1126, finish_join.c:48-141): one buffer per rank
 *     (num_rank_allocated of them), each 64 DPU blocks of the rank's largest
 *     block size; a block holds the pairs whose S tuple's global partition
 *     glb_partition_hash(key) & (64 ranks - 1) is that DPU, then (0, 0)
 *     pairs; byte counts include the padding.  Plain plan only (bank_set =
 *     rank_set = 1, join_alg 0, 2 or 3), ranks a power of two <= 64; heavy-
 *     key handling is off in this layout.  (0, 0) is also a legal pair when
 *     rid 0 occurs on both sides, which is why it is not the default. */
enum { SPJ_RESULT_EXACT = 0, SPJ_RESULT_RANK_PADDED = 1 };
int spj_set_result_layout(spj_instance_t *inst, int layout);
/* FinishJoin in either layout: up to cap buffers (library-allocated, freed
 * with spj_free_result) and their byte counts; *nbufs receives the number. */
int spj_finish_buffers(spj_instance_t *inst, spj_pair_t **bufs, uint64_t *bytes, int cap, int *nbufs);
/* ExecuteJoin("phj" | "smj", design), query.cc:199: start + finish. */
int spj_execute(spj_instance_t *inst, const char *join_type, const spj_join_design_t *design, spj_pair_t **pairs,
                uint64_t *count);
int spj_last_phase_times(spj_instance_t *inst, spj_phase_times_t *times);
/* Join plan, replacing IDPHandler::SetQueryTree (pipeline.cc:251-279): a plan
 * document {"query_tree": [{"operator": ...}, ...]} with the operators of
 * execution.hpp's join_json / bg_join_json (SEND_DATA_OPT left/right,
 * SYNC_THREADS, GLB_PARTITION, BG_BROADCAST + CHIPWISE_BROADCAST,
 * GLB_CHIPWISE_PARTITION, ROTATE_AND_STREAM, RNS_JOIN / IDP_JOIN,
 * RECV_DATA_OPT).  It is validated (each relation sent, then partitioned
 * once; one join after both partitions; the result received after it) and
 * mapped onto the GPU phases: BG_BROADCAST turns heavy-key handling on.
 * Other operators are rejected (SPJ_EINVAL).  NULL returns to the design's
 * choice at StartJoin (bank_set or rank_set > 1: the BG_BROADCAST plan). */
int spj_set_query_plan(spj_instance_t *inst, const char *plan_json);
/* Validate a plan without an instance: node count and whether it replicates. */
int spj_plan_check(const char *plan_json, int *num_nodes, int *replicate);
/* RecordLogs(start, rank_group), queryHostFunc.cc:183-241: the last join's
 * timeline, one line per plan node
 *   "rank_group-0-index-start_ms-stop_ms-OPERATOR-inputs"
 * with times from the CLOCK_MONOTONIC instant `start` (each node spans the
 * GPU phase doing its work; R's and S's partitions share an interval when
 * one launch scattered both), then "rnc_xfer_bytes B rnc_num_packets P" for
 * the partitioned R and S (B bytes moved as P 8-byte tuples).  *len receives
 * the text length; buf (cap > length) receives the text, or may be NULL. */
int spj_timeline(spj_instance_t *inst, int64_t start_sec, int64_t start_nsec, int rank_group, char *buf,
                 uint64_t cap, uint64_t *len);
void spj_free_result(spj_pair_t *pairs);
void spj_destroy(spj_instance_t *inst);

/* ---- serving: resident tables, budgeted admission, continuous batching ----
 * No counterpart in the reference.  Tables are radix-partitioned once and
 * stay in HBM; submitted probe batches (S) are admitted in arrival order
 * while their HBM footprint fits the budget and packed into batches that
 * run on `slots` concurrent streams. */
typedef struct spj_server spj_server_t;
spj_server_t *spj_server_create(uint64_t hbm_budget_bytes, int slots);
int spj_server_add_table(spj_server_t *sv, const spj_tuple_t *R, uint64_t nR, int *table_id);
/* S stays borrowed until spj_server_wait returns for this request. */
int spj_server_submit(spj_server_t *sv, int table_id, const spj_tuple_t *S, uint64_t nS, uint64_t *request_id);
/* Blocks until the request is done; pairs are freed with spj_free_result;
 * latency_ms = completion minus submission. */
int spj_server_wait(spj_server_t *sv, uint64_t request_id, spj_pair_t **pairs, uint64_t *count, double *latency_ms);
int spj_server_stats(spj_server_t *sv, uint64_t *batches, uint64_t *requests);
void spj_server_destroy(spj_server_t *sv);

/* ---- autotuner: per size / skew bucket parameter table ---- */
/* Set / get the parameters of the bucket (ceil log2 nR, ceil log2 nS/nR,
 * skew class of zipf) -- used by every later call whose config fields are -1. */
int spj_tuner_set(uint64_t nR, uint64_t nS, double zipf, const spj_config_t *cfg);
int spj_tuner_get(uint64_t nR, uint64_t nS, double zipf, spj_config_t *cfg);
int spj_tuner_clear(void);
int spj_tuner_entries(void);
int spj_tuner_save(const char *path);
int spj_tuner_load(const char *path);
/* Time candidate radix bits / items-per-thread / LDS table sizes on these
 * device relations and store the fastest for their bucket. */
int spj_autotune(const spj_tuple_t *R, uint64_t nR, const spj_tuple_t *S, uint64_t nS, spj_pair_t *out, double zipf,
                 int reps, float *best_ms);

/* ---- synthetic workloads (seeded; identical bytes for a given seed) ---- */
/* R: keys = random permutation of 1..nR, rid = row. */
int spj_gen_r(spj_tuple_t *R, uint64_t nR, uint64_t seed);
/* S: keys drawn Zipf(zipf) over R's keys in a seeded random rank order
 * (zipf = 0: uniform), rid = row.  threads <= 0: all cores. */
int spj_gen_s(spj_tuple_t *S, uint64_t nS, const spj_tuple_t *R, uint64_t nR, double zipf, uint64_t seed,
              int threads);

#ifdef __cplusplus
}
#endif
#endif
