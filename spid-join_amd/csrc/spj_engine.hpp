// spj_engine.hpp -- host engine: plans a batch of join instances, owns the
// device work buffers of one execution context (one HIP stream), launches
// the phase sequence and reads back exact result counts.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "spj_common.hpp"

namespace spj {

// Per-instance parameters after the tuner / defaults have been applied.
struct Resolved {
    uint32_t bits = 0;        // radix bits of the partition pass
    uint32_t table_log2 = 14; // LDS hash-table slots (log2)
    uint32_t alg = 0;         // 0 hash (+NL for tiny), 2 sort-merge, 3 nested loop
    uint32_t heavy = 0;       // heavy-key detection and direct emission
    uint32_t heavy_min = 16;  // summed sample count for a key to be heavy
    uint32_t ipt = 12;        // scatter tuples per thread (round = 1024 * ipt)
    uint32_t tile = 65536;    // tuples per tile of the partition pass
    uint32_t smax = 131072;   // S tuples per work item
    uint32_t cells = 2;       // S layout: 0 counted (histogram pass), 1 cells (none), 2 auto
    uint32_t cell_pct = 0;    // cell capacity in % of the mean cell fill (0 = automatic)
    uint32_t cell_mean_x2 = 0;// automatic capacity from twice the mean fill (a second radix pass)
    uint32_t passes = 1;      // radix partition passes: 1, or 2 (bits is the first pass's)
};

struct InstSpec {
    const uint2 *R = nullptr;
    uint32_t nR = 0;
    const uint2 *S = nullptr;
    uint32_t nS = 0;
    uint2 *out = nullptr;     // capacity >= nS pairs
    Resolved cfg;
    // Optional: R already radix-partitioned with cfg.bits (a resident table):
    // Rp_pre holds the partitioned tuples and pst_pre (device) the F + 1
    // partition starts.  Only S is partitioned then; heavy handling is off.
    const uint2 *Rp_pre = nullptr;
    const uint32_t *pst_pre = nullptr;
    // R must end up in contiguous partitions (export_partitioned_r); S too
    bool r_contiguous = false, s_contiguous = false;
    // digit = (SuperFastHash(key) >> shift) & (2^cfg.bits - 1): a second
    // radix pass over one first-pass partition takes the next bits
    uint32_t shift = 0;
    // partition R and S only (heavy pairs emitted to the back of out), no
    // join: the first of two passes.  A batch is all or none.
    bool partition_only = false;
    // Optional: R (S) as segments (offset, length) of R (S), the relation
    // being their concatenation and nR (nS) their total -- a first-pass
    // partition that lies in cells.  No heavy keys and no no-partition join
    // for such an instance.
    std::vector<uint2> rseg, sseg;
};

// Milliseconds per phase (HIP events on the context's stream).
struct PhaseTimes {
    float heavy = 0, hist = 0, scan = 0, scatter = 0, items = 0, join = 0, total = 0;
    float scatter_r = -1;               // R's own scatter launch (ran first); -1: R and S in one launch
    uint32_t heavy_keys = 0;            // instance 0
    uint64_t r_parted = 0, s_parted = 0;  // instance 0: tuples written to partitions
    uint32_t s_cells = 0, r_cells = 0;    // instance 0: S / R layout was cells
    uint32_t passes = 1;                  // radix partition passes
};

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes);
    void release();
    template <class T> T *as() const { return reinterpret_cast<T *>(p); }
};

// Page-locked host memory (hipHostMalloc), grown on demand: staging for the
// descriptor upload and the control-word read-back, so both copies are
// asynchronous DMA instead of staged pageable copies.
struct PinBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes);
    void release();
    template <class T> T *as() const { return reinterpret_cast<T *>(p); }
};

struct DeviceInfo {
    int device = 0;
    int cus = 0;
    size_t lds_per_block = 0;
    std::string arch;
};

// Query (once) the current device.
const DeviceInfo &device_info();

class Context {
public:
    Context();
    ~Context();
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;

    hipStream_t stream() const { return stream_; }

    // Run a batch.  counts/errs receive per-instance results.  When timed is
    // non-null, events bracket every phase.  Returns hipSuccess or the first
    // HIP error; join-level errors are in errs.
    hipError_t run(const std::vector<InstSpec> &specs, std::vector<uint32_t> &counts, std::vector<uint32_t> &errs,
                   PhaseTimes *timed);

    // Enqueue only (no host synchronisation): results land in the device
    // arrays returned by out_counts()/errs() when the stream reaches them.
    hipError_t enqueue(const std::vector<InstSpec> &specs, PhaseTimes *timed);
    hipError_t collect(size_t ninst, std::vector<uint32_t> &counts, std::vector<uint32_t> &errs, PhaseTimes *timed);

    // Launch sequences replayed as HIP graphs: graphs captured, replays.
    uint64_t graph_captures() const { return graph_captures_; }
    uint64_t graph_replays() const { return graph_replays_; }

    // Record, for every join work item, (instance, partition, first output
    // position, pairs written); read them with item_extents() after collect().
    // Positions are before the compaction of holes (see collect).
    void set_record_extents(bool on);
    hipError_t item_extents(std::vector<uint4> &ext);

    // After run()/collect() of a batch, copy instance i's partitioned R and
    // its F + 1 partition starts to caller-owned device buffers.
    hipError_t export_partitioned_r(size_t inst, uint2 *rp_dst, uint32_t *pstart_dst);

    // After run()/collect() of a partition_only batch: instance i's
    // partitioned R and S (in this context's buffers, valid until its next
    // run) as segments (offset into rp / sp, length), read back to the host:
    // partition p of R is segs_r[p * ng_r + k], k < ng_r -- one segment per
    // partition in the counted layout (ng = 1), one per scatter group in the
    // cell layout (ng = 8).  The same for S.
    hipError_t partition_segments(size_t inst, const uint2 **rp, const uint2 **sp, uint32_t &ng_r,
                                  std::vector<uint2> &segs_r, uint32_t &ng_s, std::vector<uint2> &segs_s);

    // Make out[0, total) dense: copy segments (src, dst, len) of out, in
    // order, through a temporary buffer (sources and destinations may
    // overlap).  On this context's stream, synchronised.
    hipError_t gather(uint2 *out, const std::vector<uint4> &segs, uint64_t total);

private:
    hipError_t enqueue_plan(const std::vector<InstSpec> &specs, PhaseTimes *timed);
    void drop_graph();
    hipStream_t stream_ = nullptr;
    // graph replay of a repeated call (see enqueue)
    bool graphs_ = true;            // off after a failed capture, or with SPJ_GRAPHS=0
    bool last_valid_ = false;       // h_specs_ is the last call, launched directly, buffers unmoved since
    bool capturing_ = false;
    hipGraphExec_t graph_exec_ = nullptr;
    uint64_t graph_captures_ = 0, graph_replays_ = 0;
    hipEvent_t ev_[8] = {};   // [0, 6]: phase boundaries; [7]: end of R's own scatter launch
    bool split_timed_ = false;  // the last timed enqueue recorded ev_[7]
    bool hist_timed_ = false;   // ... ev_[2] (it launched a histogram pass)
    bool partition_only_ = false;  // the last enqueue partitioned only (no join)
    bool scan_timed_ = false;   // ... ev_[3] (it launched a scan)
    // desc_: every descriptor array of a batch (instances, tiles, scan
    // segments and blocks, heavy jobs) in one block, uploaded by one copy
    DevBuf desc_, counts_, pstart_;
    DevBuf rp_, sp_, items_, ctrl_, heavy_, cand_, holes_, tmp_, segbuf_, tiles_, nphj_, ext_;
    bool record_ext_ = false;
    PinBuf pin_desc_, pin_ctrl_;
    hipEvent_t desc_ev_ = nullptr;   // the last descriptor upload has left pin_desc_
    bool desc_pending_ = false;
    std::vector<HeavyJob> h_jobs_;
    std::vector<NphjJob> h_nphj_;   // no-partition hash joins of the batch
    uint32_t max_holes_ = 0;
    size_t specs_n_ = 0;
    std::vector<uint32_t> h_ns_;
    hipError_t compact(size_t inst, uint32_t lo, uint32_t hi, uint32_t fill, const std::vector<uint4> &holes);
    std::vector<InstDesc> h_insts_;
    std::vector<TileRun> h_runs_;   // expanded into tiles_ on the device
    uint32_t ntiles_ = 0;
    std::vector<CellGroup> h_cells_;
    // the batch of the last enqueue, redone with counted partitions when an
    // S cell overflowed (force_counted_ set meanwhile)
    std::vector<InstSpec> h_specs_;
    bool force_counted_ = false;
    std::vector<ScanSeg> h_segs_;
    std::vector<uint32_t> h_ctrl_;
    uint32_t max_bits_ = 0, max_table_ = 0;
    int max_ipt_ = 8;
    bool any_heavy_ = false;
};

// Default parameters for an instance of the given sizes (no tuner entry).
Resolved default_params(uint32_t nR, uint32_t nS);
// Default partition-pass tile size for a relation pair whose larger side has n tuples.
uint32_t default_tile(uint64_t n);

}  // namespace spj
