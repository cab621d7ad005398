// spj_kernels.hpp -- host-side launchers for the kernels in spj_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "spj_common.hpp"

namespace spj {

size_t lds_hist(uint32_t max_bits);
size_t lds_scatter(uint32_t max_bits, int ipt);
size_t lds_join();
uint32_t join_table_cap();
uint32_t cand_words();   // words of one job's candidate map (keys, counts, found flag)

hipError_t launch_hist(const InstDesc *insts, const TileDesc *tiles, uint32_t ntiles, uint32_t *counts,
                       uint32_t max_bits, hipStream_t st);
hipError_t launch_scan(const ScanSeg *segs, uint32_t nsegs, uint32_t *counts, uint32_t *pstart, uint32_t max_bits,
                       hipStream_t st);
// Persistent scatter over `grid` workgroups (rounded up to a multiple of 8);
// counts holds the group cursors (scanned, or cell bases) and is advanced in
// place; tq holds 8 zeroed tile-queue counters (one per XCD).
hipError_t launch_scatter(const InstDesc *insts, const TileDesc *tiles, uint32_t ntiles, uint32_t *counts,
                          uint32_t *out_hi, uint32_t *err, uint32_t *tq, uint32_t max_bits, int ipt, uint32_t grid,
                          hipStream_t st);
// Sample blocks for heavy detection of an S of nS tuples (at most
// max_blocks, evenly strided); *stride receives the stride in blocks.
uint32_t heavy_sample_blocks(uint32_t nS, uint32_t max_blocks, uint32_t *stride);
// Heavy-key detection for every job at once: a sampling launch over every
// job's pre-sample blocks (blocks[0] in all), a second one over the other
// sample blocks (blocks[1]; a job skips them when its pre-sample found no
// candidate), then one selection workgroup per job.  Each job owns
// cand_words() zeroed words at cand_base + cand_off.  The sampling launches
// have at most `grid` workgroups (one per CU).
hipError_t launch_heavy_detect(const InstDesc *insts, const HeavyJob *jobs, uint32_t njobs, const uint32_t blocks[2],
                               uint32_t *cand_base, uint32_t *nheavy, uint32_t grid, hipStream_t st);
// One launch before a batch's passes: zero the control words, heavy tables,
// candidate maps and counted relations' counters, expand the tile runs, set
// the cell cursors.
hipError_t launch_prep(const PrepArgs &a, hipStream_t st);
// Work items of every instance, with their partitions' position maps.
// spart[i] and rpart[i] receive instance i's partitioned S and R tuple counts.
hipError_t launch_make_items(const InstDesc *insts, uint32_t ninst, const uint32_t *pstart, const CellGroup *cells,
                             const uint32_t *counts, uint32_t smax, WorkItem *items, uint32_t *nitems,
                             uint32_t *spart, uint32_t *rpart, hipStream_t st);
// ext (optional, one entry per work item): (instance, partition, first
// output position, pairs written) of every item.  range: some instance is a
// sort-merge join (the kernel with key-range buckets is launched).
hipError_t launch_join(const InstDesc *insts, const WorkItem *items, uint32_t *ctrl,
                       uint32_t *out_lo, uint32_t *out_fill, uint32_t *err, uint4 *holes, uint32_t max_holes,
                       uint4 *ext, uint32_t grid, bool range, hipStream_t st);
// No-partition hash join of every job's instance: zero the tables, insert
// R (a duplicate key flags kErrDuplicateKey), then probe S and append each
// match at out_lo[inst] (out_fill advanced by the same count).
hipError_t launch_nphj(const InstDesc *insts, const NphjJob *jobs, uint32_t njobs, uint64_t *tables,
                       uint64_t table_words, uint32_t build_blocks, uint32_t probe_blocks, uint32_t *out_lo,
                       uint32_t *out_fill, uint32_t *err, bool clear, hipStream_t st);
// The no-allocation join (no partition pass) of one instance of at most
// kNoAllocMaxR R tuples: descriptors, control words and table live in a
// static device array (one per device, held until the process ends), so
// nothing is allocated.  ctrl_out receives (pairs reserved, pairs written,
// error flags).  The call waits for the stream.  Calls share the device's
// array: they must not overlap on one device.
constexpr uint32_t kNoAllocMaxR = 1u << 20;
constexpr uint32_t kNoAllocMaxDevices = 64;
hipError_t run_nphj_noalloc(const uint2 *R, uint32_t nR, const uint2 *S, uint32_t nS, uint2 *out, uint32_t ctrl_out[3],
                            hipStream_t st);
// Look up (once per device) the static array's address on the current device.
hipError_t scratch_on_device(void **ptr);
hipError_t launch_pairs_checksum(const uint2 *p, uint64_t n, unsigned long long *acc, hipStream_t st);
hipError_t launch_gather_segments(const uint2 *src, uint2 *dst, const uint4 *segs, uint32_t nsegs, hipStream_t st);

}  // namespace spj
