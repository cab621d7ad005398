// spj_common.hpp -- descriptors shared by the host engine and the HIP kernels.
//
// A "batch" is a set of independent join instances.  A single join is a batch
// of one; the batched workload (thousands of uneven joins) and the serving
// loop (many probe batches against resident tables) use the same kernels.
//
// Layout in HBM (all 8-byte AoS tuples {uint32 key; uint32 rid}, the
// reference's kv_pair_t, type_defs.hpp:53-56):
//   R, S                 caller's inputs (read only)
//   Rp, Sp               radix-partitioned copies, one region per instance,
//                        partition p of an instance occupying
//                        [pstart[p], pstart[p+1]) of its region
//   counts               per (instance, relation, scatter group) digit
//                        counts, scanned in place into the group's start
//                        offsets, then advanced as write cursors
//   cells                (large relations, no histogram pass) group g's
//                        tuples of digit d go to a cell of fixed capacity
//                        cap_g at cell0_g + d * cap_g of the region; the
//                        cursor starts at the cell's base and ends at its
//                        fill
//   out                  (R rid, S rid) pairs, one region per instance with
//                        capacity |S_i|; exact count in out_count[i]
#pragma once
#include <stdint.h>

namespace spj {

constexpr uint32_t kWave = 64;
constexpr uint32_t kEmptyKey = 0u;          // key 0 = empty / skipped (glb_partition_count.c:192)
constexpr uint32_t kHeavySlots = 2048;      // heavy (key -> R rid) table, LDS resident
// One instance's heavy table in device memory (and in LDS, in the same
// order): kHeavySlots keys (0 = empty), kHeavySlots R rids, then a presence bitmap (bit s set: slot s's key has an
// R tuple and its rid is valid).  Presence is kept apart from the rid because
// every 32-bit value is a legal rid (kv_pair_t.rvalue is a full uint32).
constexpr uint32_t kHeavyPresWords = kHeavySlots / 32;
constexpr uint32_t kHeavyWords = 2 * kHeavySlots + kHeavyPresWords;
constexpr uint32_t kHeavyMaxKeys = 1024;    // at most half full
constexpr uint32_t kMaxBits = 12;           // at most 4096 partitions per pass
constexpr uint32_t kMaxDigitBits = 19;      // shift + bits of a pass: digits come from SuperFastHash bits 0..18
constexpr uint32_t kScatterNT = 1024;       // scatter workgroup size
constexpr uint32_t kScatterGroups = 8;      // scatter groups (one per XCD) per relation
constexpr uint32_t kSinkTuples = kScatterNT * 12;   // largest scatter round

// error bits in err[inst]
constexpr uint32_t kErrDuplicateKey = 1u;   // a2a_srj_build.c:265-271 (VALIDATION)
constexpr uint32_t kErrOutOverflow = 4u;    // more pairs than the output capacity
constexpr uint32_t kErrCellOverflow = 8u;   // an S cell received more tuples than its capacity

struct InstDesc {
    const uint2 *R;          // input R (device)
    const uint2 *S;          // input S (device)
    uint2 *Rp;               // partitioned R region
    uint2 *Sp;               // partitioned S region
    uint2 *out;              // output pairs region (capacity nS)
    uint32_t nR, nS;
    uint32_t bits;           // radix bits (F = 1 << bits)
    uint32_t heavy;          // 1: heavy table in use for this instance
    uint64_t pstartR;        // offset (in uint32) of R partition starts (F+1 entries)
    uint64_t pstartS;        // offset of S partition starts (F+1 entries)
    uint32_t *heavy_keys;    // kHeavySlots keys (0 = empty)
    uint32_t *heavy_rids;    // kHeavySlots R rids, then kHeavyPresWords presence bits
    const uint32_t *heavy_count;   // number of heavy keys found (device)
    uint32_t table_log2;     // LDS hash table slots (log2) for this instance
    uint32_t alg;            // 0 hash (auto NL for tiny), 2 sort-merge, 3 nested loop
    // Cell layout of S (s_) and R (r_): 1 + index of the relation's first
    // of kScatterGroups CellGroups (0 = counted layout), and the region
    // offset of a kSinkTuples sink for runs that overflow a cell.
    uint32_t s_cells, s_sink;
    uint32_t r_cells, r_sink;
    // The digit of a key is (SuperFastHash(key) >> shift) & (2^bits - 1):
    // shift 0 for a one-pass partition and for the first of two passes; the
    // second pass partitions each first-pass partition by the next bits.
    uint32_t shift;
    uint32_t pad_;
};

// Cell layout of one scatter group of one relation: the group's cursors
// at cnt_off (F of them), digit d's cell at cell0 + d * cap of the Sp
// region (cap = 0: the group holds no S tile of the instance).
struct CellGroup {
    uint64_t cnt_off;
    uint32_t cell0;
    uint32_t cap;
    uint32_t bits;
    uint32_t pad;
};

// Where a partition's tuples lie in the cell layout: partition-relative
// position v is at region[v + D[g]] for the largest g with V[g] <= v
// (V[0] = 0; V[g] = total of cells 0..g-1).  The contiguous layout is
// V = {0, ~0, ...}, D = {0}.
struct SMap {
    uint32_t V[kScatterGroups];
    uint32_t D[kScatterGroups];
};

// One tile of one relation of one instance.  cnt_off is the offset of the
// F digit counters of the tile's scatter group: histogram counts before the
// scan, write cursors after it (cell layout: cursors from the start).
struct TileDesc {
    uint32_t inst;
    uint32_t rel;            // 0 = R, 1 = S
    uint32_t begin, end;     // tuple range inside the relation
    uint64_t cnt_off;
    uint32_t cell0, cap;     // cell layout: the group's cells (cap = 0: counted layout)
};

// The tiles of one relation of one instance, or of one segment of it: ntiles
// tiles of `tile` tuples (the last one shorter) at tile-list positions
// [first, first + ntiles),
// in the scatter launch of tiles [launch0, launch0 + launch_n), whose
// position decides a tile's scatter group.  The relation's group-0 counters
// are at cnt_off0; cells = 1 + index of its first CellGroup (0: counted).
// The device expands runs into TileDescs.
struct TileRun {
    uint32_t inst, rel;
    uint32_t first, ntiles;
    uint32_t tile, len;
    uint64_t cnt_off0;
    uint32_t cells, bits;
    uint32_t launch0, launch_n;
    uint32_t base, pad_;     // the run covers tuples [base, base + len) of the relation
};

// Scatter group of position p in a scatter launch: XCD x takes tiles x,
// x + 8, x + 16, ... (dealt round-robin, so that every XCD gets the same
// mix of relations and instances and the XCDs finish together).
constexpr uint32_t tile_group(uint32_t p) { return p % kScatterGroups; }

// Everything k_prep does for a batch (see there): words to zero, the tile
// runs to expand, the cell groups whose cursors to set.
struct PrepArgs {
    uint32_t *ctrl;
    uint64_t nctrl;
    uint32_t *heavy;
    uint64_t nheavy;
    uint32_t *cand;
    uint64_t ncand;
    uint32_t *counts;
    uint64_t nzero;           // counters of counted relations: counts[0, nzero)
    const TileRun *runs;
    uint32_t nruns;
    uint32_t ntiles;
    TileDesc *tiles;
    const CellGroup *cells;
    uint32_t ncells;
    uint32_t pad;
    unsigned long long *nphj;   // the no-partition joins' tables: nphj[0, nnphj) cleared
    uint64_t nnphj;
};

// One (instance, relation) whose group counts are scanned together: t_count
// groups (kScatterGroups), group g's counters at cnt_off0 + g * F;
// partition starts at pstart_off (F + 1 entries).
struct ScanSeg {
    uint32_t inst;
    uint32_t rel;
    uint32_t bits;
    uint32_t t_count;
    uint64_t cnt_off0;
    uint64_t pstart_off;
};

// Heavy-key detection of one instance: nblocks sample blocks, every
// stride-th 8K-tuple block of S.  Phase 0 samples every kPreSample-th of
// them; phase 1 samples the rest, and only if phase 0 found a candidate.
// first_block[p] is the job's first block in phase p's launch.
struct HeavyJob {
    uint32_t inst;
    uint32_t stride;
    uint32_t first_block[2];
    uint32_t nblocks;
    uint32_t min_count;
    uint64_t cand_off;       // candidate map: keys, counts, then the found flag
};
constexpr uint32_t kPreSample = 16;

// One no-partition hash join (join_alg 1): instance `inst`'s R in an open-
// addressing table of 2^log2cap 8-byte slots at tables + table.  Its blocks
// in the build and probe launches start at first_build and first_probe
// (kNphjBuildTuples R and kNphjProbeTuples S tuples per block).
struct NphjJob {
    uint32_t inst;
    uint32_t log2cap;
    uint64_t table;
    uint32_t first_build;
    uint32_t first_probe;
};
constexpr uint32_t kNphjPPT = 16;                       // S tuples per thread (256-thread blocks)
constexpr uint32_t kNphjProbeTuples = 256 * kNphjPPT;
constexpr uint32_t kNphjBuildTuples = 256 * 4;

// One join work item: partition `part` of instance `inst`, S range
// [sb, se) and the partition's R range [rb, re), absolute positions in the
// instance's partitioned regions (so the join reads one descriptor, not the
// partition starts as well).  The S and R positions go through the item's
// position maps (the identity for counted relations; with a cell layout the
// ranges are partition-relative).  40 words, all the join reads of an item.
struct WorkItem {
    uint32_t inst;
    uint32_t part;
    uint32_t sb, se;
    uint32_t rb, re;
    uint32_t pad[2];
    SMap smap, rmap;
};
constexpr uint32_t kItemWords = sizeof(WorkItem) / 4;

}  // namespace spj
