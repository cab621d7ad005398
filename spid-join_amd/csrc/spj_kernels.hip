// spj_kernels.hip -- CDNA4 (gfx950) kernels for the skew-resistant equi-join.
//
// Phase map (reference operator -> kernel):
//   GLB_PARTITION_COUNT (glb_partition_count.c)          -> k_hist
//   CALCULATE_PAGE_HISTOGRAM (queryHostFunc.cc:875-960)   -> k_scan
//   GLB_PARTITION_PACKET + ROTATE_AND_STREAM              -> k_scatter
//   BG_BROADCAST / CHIPWISE_BROADCAST (SPID replication)  -> k_heavy_sample,
//        k_heavy_select, heavy capture in k_hist (R) and
//        direct heavy emission in k_scatter (S)
//   PHJ_BUILD + PHJ_PROBE (a2a_srj_build.c,
//        a2a_srj_probe_inner_mt.c)                         -> k_join (hash)
//   MPSM_JOIN_* (mpsm_join_*.c)                           -> k_join (sort-merge)
//   NESTED_LOOP_JOIN (nested_loop_join.c)                 -> k_join (nested loop)
//
// All work is integer: hashing, LDS atomics, prefix sums, coalesced copies.
// The bound is HBM bandwidth; no MFMA.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <vector>

#include "spj_common.hpp"
#include "spj_kernels.hpp"

namespace spj {

// ---------------------------------------------------------------------------
// Accesses through the device pointers held in InstDesc.  Those pointers are
// generic, and generic (FLAT) instructions count on lgkmcnt as well as on
// vmcnt, so every LDS wait and barrier would also wait for the streaming
// loads and stores in flight.  These helpers access them as global memory.
// ---------------------------------------------------------------------------
typedef __attribute__((address_space(1))) uint64_t gu64_t;
typedef __attribute__((address_space(1))) uint32_t gu32_t;

__device__ __forceinline__ uint2 gld(const uint2 *p, uint64_t i)
{
    const uint64_t x = reinterpret_cast<const gu64_t *>(reinterpret_cast<uintptr_t>(p))[i];
    return make_uint2((uint32_t)x, (uint32_t)(x >> 32));
}

__device__ __forceinline__ void gst(uint2 *p, uint64_t i, uint2 v)
{
    reinterpret_cast<gu64_t *>(reinterpret_cast<uintptr_t>(p))[i] = (uint64_t)v.x | ((uint64_t)v.y << 32);
}

__device__ __forceinline__ uint32_t gld32(const uint32_t *p, uint64_t i)
{
    return reinterpret_cast<const gu32_t *>(reinterpret_cast<uintptr_t>(p))[i];
}

__device__ __forceinline__ void gst32(uint32_t *p, uint64_t i, uint32_t v)
{
    reinterpret_cast<gu32_t *>(reinterpret_cast<uintptr_t>(p))[i] = v;
}

// ---------------------------------------------------------------------------
// Hashes (device).  Same functions as the reference's src/dpu/src/hash.c.
// ---------------------------------------------------------------------------

// Paul Hsieh's SuperFastHash of the 4 little-endian key bytes, initial hash
// = length 4 (hash.c:117-166).  Its low bits select the global partition.
__device__ __forceinline__ uint32_t sfh(uint32_t key)
{
    uint32_t h = 4u + (key & 0xFFFFu);
    uint32_t tmp = ((key >> 16) << 11) ^ h;
    h = (h << 16) ^ tmp;
    h += h >> 11;
    h ^= h << 3;
    h += h >> 5;
    h ^= h << 4;
    h += h >> 17;
    h ^= h << 25;
    h += h >> 6;
    return h;
}

// The low bits of sfh(key) alone, for digits of up to 19 bits: the step
// h ^= h << 25 changes only bits 25 and up, and the final h += h >> 6 brings
// bits up to 24 down no further than bit 18, so bits 0..18 of the result do
// not depend on it.
__device__ __forceinline__ uint32_t sfh_low(uint32_t key)
{
    uint32_t h = 4u + (key & 0xFFFFu);
    uint32_t tmp = ((key >> 16) << 11) ^ h;
    h = (h << 16) ^ tmp;
    h += h >> 11;
    h ^= h << 3;
    h += h >> 5;
    h ^= h << 4;
    h += h >> 17;
    h += h >> 6;
    return h;
}
static_assert(kMaxBits <= 19 && kMaxDigitBits <= 19, "sfh_low gives digits of up to 19 bits");

// Bucket of a key in the join's LDS table: Fibonacci (multiply-shift)
// hashing of the key.  Partitions are chosen by SuperFastHash bits, so the
// two are independent.
__device__ __forceinline__ uint32_t bucket_of(uint32_t key, uint32_t log2_buckets)
{
    return (key * 0x9E3779B1u) >> (32 - log2_buckets);
}

// Heavy table: kHeavySlots / 4 buckets of 4 slots; a key lives in one of
// two buckets (2-choice), so a lookup is two independent 16-byte LDS reads
// and never loops.  At most kHeavyMaxKeys keys (half full): placing 1024
// keys hottest first into 512 such buckets leaves about one key in 100
// selections without a slot (it stays an ordinary key).
constexpr uint32_t kHeavyBucketBits = 9;    // 512 buckets x 4 slots = kHeavySlots
static_assert(kHeavySlots == (4u << kHeavyBucketBits), "heavy table: 4-slot buckets");
// The two buckets of a key come from its SuperFastHash h = sfh(key), which
// the scatter computes for the digit anyway: bits 13..21 and 23..31, above
// the digit bits of any pass that uses the heavy table (radix bits <= 12).
// Two bit fields cost two VALU operations where multiplicative hashes cost
// two quarter-rate 32-bit multiplies and more.
static_assert(kMaxBits <= 13 && kHeavyBucketBits <= 9, "heavy buckets: sfh bits 13..21 and 23..31");
constexpr int kHeavyLookupGroup = 4;   // scatter: tuples whose heavy lookups are issued together
__device__ __forceinline__ uint32_t heavy_b1(uint32_t h) { return (h >> 13) & ((1u << kHeavyBucketBits) - 1u); }
__device__ __forceinline__ uint32_t heavy_b2(uint32_t h) { return h >> (32 - kHeavyBucketBits); }

// Sub-table of a partition that needs Q > 1 LDS tables: SFH bits above the
// digit bits of every pass (bits = shift + radix bits), so they do not
// depend on the bits that chose the partition.
__device__ __forceinline__ uint32_t sub_of(uint32_t h, uint32_t bits, uint32_t Q)
{
    return __umulhi((h >> bits) * 0x9E3779B1u, Q);
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t mbcnt(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Inclusive prefix sum over the 64 lanes of a wave, by DPP: shifts by 1, 2,
// 4 and 8 lanes inside each row of 16 (lanes shifted in from outside the row
// read 0), then lane 15 of row 0 added to row 1 and of row 2 to row 3
// (row_bcast:15), then lane 31 added to rows 2 and 3 (row_bcast:31).
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t x)
{
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, true);   // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, true);   // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, true);   // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, true);   // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);  // row_bcast:15, rows 1 and 3
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);  // row_bcast:31, rows 2 and 3
    return x;
}

// In-place exclusive scan of a[0..n) in LDS by the whole workgroup; returns
// the total.  wtmp needs blockDim/64 + 1 words.
__device__ uint32_t block_exclusive_scan(uint32_t *a, uint32_t n, uint32_t *wtmp)
{
    // The thread index is made opaque, so that the lane masks and shuffle
    // addresses derived from it are formed at each call: hoisted out of a
    // caller's loop they would hold registers for the whole loop (and spill
    // in the scatter, whose reloads wait behind the round's stores).
    uint32_t tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const uint32_t nt = blockDim.x;
    const uint32_t nw = nt / kWave, wid = tid / kWave, lane = tid & (kWave - 1);
    uint32_t per = (n + nt - 1) / nt;
    uint32_t b = tid * per, e = min(n, b + per);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; i++) s += a[i];
    const uint32_t x = wave_incl_sum(s);
    if (lane == kWave - 1) wtmp[wid] = x;
    __syncthreads();
    if (tid < kWave) {
        uint32_t v = (tid < nw) ? wtmp[tid] : 0u;
        const uint32_t z = wave_incl_sum(v);
        if (tid < nw) wtmp[tid] = z - v;
        if (tid == kWave - 1) wtmp[nw] = z;
    }
    __syncthreads();
    uint32_t ex = wtmp[wid] + x - s;
    for (uint32_t i = b; i < e; i++) { uint32_t v = a[i]; a[i] = ex; ex += v; }
    uint32_t total = wtmp[nw];
    __syncthreads();
    return total;
}

// Tile index of block b when T tiles are spread over 8 XCDs: the hardware
// hands block b to XCD b % 8, so XCD x gets one contiguous run of tiles and
// neighbouring partition runs are written through the same L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t T)
{
    uint32_t x = b & 7u, j = b >> 3, q = T >> 3, r = T & 7u;
    return x * q + min(x, r) + j;
}

// Number of heavy keys of an instance, read through the scalar (constant)
// cache: the count is written by an earlier kernel and only read here, and a
// scalar load does not wait behind the vector loads and stores in flight
// (a vector load would, on the shared vmcnt counter).
__device__ __forceinline__ uint32_t heavy_keys_of(const InstDesc &I)
{
    typedef const __attribute__((address_space(4))) uint32_t const_u32;
    return *reinterpret_cast<const_u32 *>(reinterpret_cast<uintptr_t>(I.heavy_count));
}

// Load the first n words of the instance's heavy table into LDS at hk: the
// keys (n = kHeavySlots), or the keys, R rids and presence bitmap, which lie
// back to back in both places (n = kHeavyWords).  One flat loop: the table
// is loaded inside the scatter's round loop, whose registers are tight.  The
// caller synchronises before using it.
__device__ __forceinline__ void load_heavy(const InstDesc &I, uint32_t *hk, uint32_t n)
{
    uint32_t i0 = threadIdx.x;   // opaque: the addresses are formed here, not held across the caller's loop
    asm volatile("" : "+v"(i0));
    for (uint32_t i = i0; i < n; i += blockDim.x) hk[i] = gld32(I.heavy_keys, i);
}

// R side: slot s's key has an R tuple with this rid.  R keys are unique, so
// each slot has one writer; the presence bit shares its word with 31 other
// slots and is set atomically.
__device__ __forceinline__ void heavy_record_rid(const InstDesc &I, uint32_t s, uint32_t rid)
{
    gst32(I.heavy_rids, s, rid);
    atomicOr(I.heavy_rids + kHeavySlots + (s >> 5), 1u << (s & 31u));
}

// Slot of key (never 0: empty slots hold key 0; h = sfh(key)) in the LDS
// heavy table, or -1 when it is not there or valid is false.  Both
// candidate buckets are read at once -- also for an invalid tuple (past the
// round's end, or the empty key, which would match an empty slot): the
// buckets are in range for any h, so a thread's lookups need no branch and
// their reads go out together.  The matching position among the 8 slots is
// chosen with constants, and the slot formed once.  hk is 16-byte aligned.
__device__ __forceinline__ int heavy_find(const uint32_t *hk, uint32_t key, uint32_t h, bool valid)
{
    const uint4 *hb = reinterpret_cast<const uint4 *>(hk);
    const uint32_t b1 = heavy_b1(h), b2 = heavy_b2(h);
    const uint4 a = hb[b1], c = hb[b2];
    uint32_t i = 8;
    i = (c.w == key) ? 7u : i;
    i = (c.z == key) ? 6u : i;
    i = (c.y == key) ? 5u : i;
    i = (c.x == key) ? 4u : i;
    i = (a.w == key) ? 3u : i;
    i = (a.z == key) ? 2u : i;
    i = (a.y == key) ? 1u : i;
    i = (a.x == key) ? 0u : i;
    const int s = (int)(4u * (i < 4u ? b1 : b2) + (i & 3u));
    return (valid && i < 8u) ? s : -1;
}

// Tile t of the list from its run (runs sorted by first, none empty).
__device__ __forceinline__ void expand_tile(const TileRun *__restrict__ runs, uint32_t nruns,
                                            const CellGroup *__restrict__ cells, TileDesc *__restrict__ tiles,
                                            uint32_t t)
{
    uint32_t lo = 0, hi = nruns;   // the last run with first <= t
    while (hi - lo > 1) {
        const uint32_t m = (lo + hi) >> 1;
        if (runs[m].first <= t) lo = m;
        else hi = m;
    }
    const TileRun r = runs[lo];
    const uint32_t g = tile_group(t - r.launch0);
    TileDesc td;
    td.inst = r.inst;
    td.rel = r.rel;
    const uint32_t b = (t - r.first) * r.tile;
    td.begin = r.base + b;
    td.end = r.base + (uint32_t)min((uint64_t)r.len, (uint64_t)b + r.tile);
    td.cnt_off = r.cnt_off0 + ((uint64_t)g << r.bits);
    td.cell0 = r.cells ? cells[r.cells - 1 + g].cell0 : 0u;
    td.cap = r.cells ? cells[r.cells - 1 + g].cap : 0u;
    tiles[t] = td;
}

// k_prep: everything a batch needs before its first pass, in one launch
// (each of these was a launch of a few microseconds plus a gap):
//   y = 0: zero the control words, the heavy tables, the candidate maps and
//          the counted relations' digit counters; expand the tile runs;
//   y = 1: the cell layout's cursors, group g's cursor of digit d at its
//          cell, cell0 + d * cap (one block per cell group).
// The counted relations' counters come before the cell relations' in the
// counter array, so the two halves never write the same words.
__global__ __launch_bounds__(256) void k_prep(PrepArgs a)
{
    if (blockIdx.y == 0) {
        const uint64_t n0 = a.nctrl, n1 = n0 + a.nheavy, n2 = n1 + a.ncand, n3 = n2 + a.nzero, n4 = n3 + a.ntiles;
        const uint64_t n5 = n4 + a.nnphj;
        for (uint64_t w = (uint64_t)blockIdx.x * 256u + threadIdx.x; w < n5; w += (uint64_t)gridDim.x * 256u) {
            if (w < n0) a.ctrl[w] = 0;
            else if (w < n1) a.heavy[w - n0] = 0;
            else if (w < n2) a.cand[w - n1] = 0;
            else if (w < n3) a.counts[w - n2] = 0;
            else if (w < n4) expand_tile(a.runs, a.nruns, a.cells, a.tiles, (uint32_t)(w - n3));
            else a.nphj[w - n4] = 0ull;
        }
    } else {
        for (uint32_t c = blockIdx.x; c < a.ncells; c += gridDim.x) {
            const CellGroup g = a.cells[c];
            const uint32_t F = 1u << g.bits;
            for (uint32_t d = threadIdx.x; d < F; d += 256u) a.counts[g.cnt_off + d] = g.cell0 + d * g.cap;
        }
    }
}

// ---------------------------------------------------------------------------
// k_hist: digit histogram of one tile (glb_partition_count.c:119-371),
// added into the counters of the tile's group (the tiles one XCD scatters).
// For R with a heavy table it also records each heavy key's R rid.  For S
// with a heavy table, heavy tuples are not counted (they never go to a
// partition).  Key 0 is skipped everywhere.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_hist(const InstDesc *__restrict__ insts, const TileDesc *__restrict__ tiles,
                                              uint32_t ntiles, uint32_t *__restrict__ counts)
{
    extern __shared__ __align__(16) uint32_t lds[];
    const uint32_t t = xcd_tile(blockIdx.x, ntiles);
    const TileDesc td = tiles[t];
    const InstDesc &I = insts[td.inst];
    const uint32_t F = 1u << I.bits, mask = F - 1u, shift = I.shift;
    uint32_t *hk = lds;                   // kHeavySlots (16-byte aligned)
    uint32_t *hist = hk + kHeavySlots;    // F
    const bool heavy = I.heavy != 0 && heavy_keys_of(I) != 0;
    for (uint32_t i = threadIdx.x; i < F; i += blockDim.x) hist[i] = 0;
    if (heavy) load_heavy(I, hk, kHeavySlots);
    __syncthreads();
    const uint2 *in = td.rel ? I.S : I.R;
    const uint32_t mode = !heavy ? 0u : (td.rel == 0 ? 1u : 2u);
    constexpr uint32_t U = 8;   // loads in flight per thread
    for (uint32_t b = td.begin; b < td.end; b += U * blockDim.x) {
        uint2 v[U];
#pragma unroll
        for (uint32_t j = 0; j < U; j++) {
            const uint32_t i = b + j * blockDim.x + threadIdx.x;
            v[j] = (i < td.end) ? gld(in, i) : make_uint2(kEmptyKey, 0u);
        }
        // heavy lookups of all U tuples first, so their LDS reads overlap
        int hs[U];
#pragma unroll
        for (uint32_t j = 0; j < U; j++) hs[j] = mode != 0 ? heavy_find(hk, v[j].x, sfh(v[j].x), v[j].x != kEmptyKey) : -1;
#pragma unroll
        for (uint32_t j = 0; j < U; j++) {
            if (v[j].x == kEmptyKey) continue;
            if (mode == 2 && hs[j] >= 0) continue;   // heavy S tuple: no partition
            atomicAdd(&hist[(sfh_low(v[j].x) >> shift) & mask], 1u);
            if (mode == 1 && hs[j] >= 0) heavy_record_rid(I, (uint32_t)hs[j], v[j].y);
        }
    }
    __syncthreads();
    // the tile's group counters (zeroed before the launch)
    uint32_t *dst = counts + td.cnt_off;
    for (uint32_t i = threadIdx.x; i < F; i += blockDim.x)
        if (hist[i]) atomicAdd(&dst[i], hist[i]);
}

// ---------------------------------------------------------------------------
// Column scan of group counts into global write offsets (the page histogram
// + cross-rank prefix of queryHostFunc.cc:329-368, 1574-1769, done on the
// device and without max-padding), one workgroup per segment (a relation of
// an instance): for each digit, an exclusive prefix down its column of
// t_count group counts (the groups' offsets inside the partition), the
// column totals scanned over the digits (the partition starts), and the
// partition start added to each group offset -- which become the scatter's
// cursors.  One launch: the segment has only kScatterGroups columns.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_scan(const ScanSeg *__restrict__ segs, uint32_t *__restrict__ counts,
                                               uint32_t *__restrict__ pstart)
{
    __shared__ uint32_t wtmp[1024 / kWave + 1];
    extern __shared__ __align__(16) uint32_t lds[];   // F digit totals, then starts
    const ScanSeg sg = segs[blockIdx.x];
    const uint32_t F = 1u << sg.bits;
    uint32_t *c = counts + sg.cnt_off0;   // group g, digit d at c[g F + d]
    for (uint32_t d = threadIdx.x; d < F; d += blockDim.x) {
        uint32_t run = 0;
        for (uint32_t g = 0; g < sg.t_count; g++) {
            const uint32_t v = c[(uint64_t)g * F + d];
            c[(uint64_t)g * F + d] = run;
            run += v;
        }
        lds[d] = run;
    }
    __syncthreads();
    const uint32_t total = block_exclusive_scan(lds, F, wtmp);   // lds[d] = start of partition d
    uint32_t *p = pstart + sg.pstart_off;
    for (uint32_t d = threadIdx.x; d < F; d += blockDim.x) {
        const uint32_t base = lds[d];
        for (uint32_t g = 0; g < sg.t_count; g++) c[(uint64_t)g * F + d] += base;
        p[d] = base;
    }
    if (threadIdx.x == 0) p[F] = total;
}

// ---------------------------------------------------------------------------
// k_scatter: persistent LDS-staged radix scatter (glb_partition_packet.c:
// 355-589 writes one 8-byte tuple per mram_write; here a round of NT*IPT
// tuples is counting-sorted by digit in LDS and written digit run by digit
// run).
// Output space comes from cursors, one per (instance, relation, group,
// digit): the scan leaves the start offset of group g's run of digit d in
// counts[cnt_off + d], and each round reserves its run with one atomicAdd
// per digit.  Group g is tiles g, g + 8, g + 16, ... of the launch, and
// block b runs on XCD b % 8 and takes that group's tiles from a queue, so
// all workgroups of one XCD share the cursors: their runs of a digit are
// adjacent, and the lines they write partly fill up in that XCD's L2
// instead of leaving half-written.  The
// order of tuples inside a partition therefore depends on timing; the
// partition contents do not.
// For S with a heavy table, heavy tuples form bucket F and are emitted
// straight to the output as (R rid, S rid).
// ---------------------------------------------------------------------------
template <int IPT>
__device__ __forceinline__ void load_tile_round(uint2 *v, const InstDesc *__restrict__ insts, const TileDesc &td,
                                                uint32_t o)
{
    constexpr uint32_t NT = kScatterNT;
    const InstDesc &I = insts[td.inst];
    // the round's base is uniform: with it in scalar registers each load
    // takes the 32-bit lane offset (global_load v, s[base]) instead of a
    // 64-bit address formed per load
    const uint2 *in = (td.rel ? I.S : I.R) + o;
    const uint32_t rn = min((uint32_t)(NT * IPT), td.end - o);
#pragma unroll
    for (int j = 0; j < IPT; j++) {
        const uint32_t li = j * NT + threadIdx.x;
        if (li < rn) v[j] = gld(in, li);
    }
}

// Write out a staged round: partition tuples stage[0, hstart) go to
// outp[bs[digit] + i] (the digit is recomputed from the key, which is
// cheaper than staging it), heavy pairs stage[hstart, total) -- whose key
// was replaced by the R rid -- go to out[hb + i - hstart].
__device__ __forceinline__ void scatter_write_out(const uint2 *stage, const uint32_t *bs, uint32_t shift,
                                                  uint32_t mask, uint32_t hstart, uint32_t total, uint2 *outp, uint2 *out,
                                                  uint32_t hb, uint32_t nS, uint32_t *err_inst)
{
    uint32_t i0 = threadIdx.x;   // opaque: the stage addresses are formed here, not held across the round loop
    asm volatile("" : "+v"(i0));
    for (uint32_t i = i0; i < hstart; i += kScatterNT) {
        const uint2 t = stage[i];
        gst(outp, bs[(sfh_low(t.x) >> shift) & mask] + i, t);
    }
    for (uint32_t i = hstart + i0; i < total; i += kScatterNT) {
        const uint32_t oi = hb + (i - hstart);
        if (oi < nS) gst(out, oi, stage[i]);
        else atomicOr(err_inst, kErrOutOverflow);
    }
}

template <int IPT>
__global__ __launch_bounds__(kScatterNT) void k_scatter(const InstDesc *__restrict__ insts,
                                                        const TileDesc *__restrict__ tiles, uint32_t ntiles,
                                                        uint32_t max_bits, uint32_t *__restrict__ counts,
                                                        uint32_t *__restrict__ out_hi, uint32_t *__restrict__ err,
                                                        uint32_t *__restrict__ tq)
{
    extern __shared__ __align__(16) uint32_t lds[];
    constexpr uint32_t NT = kScatterNT, ROUND = NT * IPT;
    constexpr uint32_t DROP = 0xFFFFFFFFu;
    constexpr uint32_t FMAX = 1u << kMaxBits;
    const uint32_t x = blockIdx.x & 7u;           // this XCD's group: tiles x + 8 k
    uint32_t *const q = tq + x;                     // its tile queue (k; zeroed before the launch)
    const uint32_t te = ntiles;
    const uint32_t FM = 1u << max_bits;
    // LDS layout
    uint32_t *hk = lds;                                             // kHeavySlots (16-byte aligned)
    uint32_t *hr = hk + kHeavySlots;                                // kHeavySlots rids + presence bits
    uint2 *stage = reinterpret_cast<uint2 *>(hr + kHeavySlots + kHeavyPresWords);   // ROUND tuples
    // two rounds' digit counters (FM + 2 each), alternating: round r counts
    // into one (counts -> starts -> write base - start) while the other
    // still holds the bases that round r - 1's write-out reads
    uint32_t *lsa = reinterpret_cast<uint32_t *>(stage + ROUND);
    uint32_t *lsb = lsa + FM + 2;
    uint32_t *wtmp = lsb + FM + 2;                                  // NT/64 + 1
    uint32_t *misc = wtmp + NT / kWave + 1;                         // [0] heavy output base, [1] queue

    // Tiles come from the XCD's queue, taken two ahead: t is being
    // scattered, tnext follows it, misc[1] holds the one after (written by
    // thread 0 a round after its atomic, read at the next tile switch).
    if (threadIdx.x == 0) {
        misc[0] = atomicAdd(q, 1u);
        misc[1] = atomicAdd(q, 1u);
        misc[2] = atomicAdd(q, 1u);
    }
    for (uint32_t i = threadIdx.x; i < 2 * (FM + 2); i += NT) lsa[i] = 0;
    __syncthreads();
    uint32_t t = x + kScatterGroups * misc[0], tnext = x + kScatterGroups * misc[1];
    if (t >= te) return;   // whole workgroup
    __syncthreads();   // misc[0] and misc[1] read before reuse
    if (threadIdx.x == 0) misc[1] = misc[2];
    bool qwant = false;   // thread 0: take the next tile from the queue in this round (its value goes to misc[1])
    TileDesc td = tiles[t];
    uint32_t o = td.begin;
    uint32_t heavy_inst = DROP;   // instance whose heavy table is in LDS
    // software pipeline: the next round's tuples (possibly of the next tile)
    // are loaded while this round is scanned, staged and written
    // positions past a round's end are not loaded: they keep an earlier
    // round's tuple, or this empty one (the heavy lookups read the table for
    // them too, and mask the result)
    uint2 nv[IPT];
#pragma unroll
    for (int j = 0; j < IPT; j++) nv[j] = make_uint2(kEmptyKey, 0u);
    load_tile_round<IPT>(nv, insts, td, o);
    __builtin_amdgcn_s_waitcnt(0);   // entering the loop with no loads in flight (see the wait below)
    bool par = false;   // which counter array this round uses
    // No barrier between a round's write-out and the next round's ranking:
    // a wave that has issued its stores goes on to rank while slower waves
    // still write, and the ranking barrier also marks the previous round's
    // write-out as done.  The next round's loads are issued after the rank
    // atomics, not at the round's start: issued while the other waves' stores
    // were still queued, they made the scatter 7% (Zipf 0) to 19% (Zipf 1.5)
    // slower on MI355X.
    for (;;) {
        uint32_t *const lst = par ? lsb : lsa;    // this round's counters
        uint32_t *const lold = par ? lsa : lsb;   // the previous round's bases
        uint32_t *const bs = lst;                 // after staging: write base - start
        uint32_t tn = t, on = o + ROUND;
        TileDesc tdn = td;
        if (on >= td.end) {
            tn = tnext;
            if (tn < te) {
                tdn = tiles[tn];
                on = tdn.begin;
            }
        }
        const bool more = tn < te;
        uint2 v[IPT];
#pragma unroll
        for (int j = 0; j < IPT; j++) v[j] = nv[j];

        const InstDesc &I = insts[td.inst];
        const uint32_t F = 1u << I.bits, mask = F - 1u, shift = I.shift;
        const bool heavy = (I.heavy != 0) && (td.rel == 1) && heavy_keys_of(I) != 0;
        // R tiles of the cell layout record each heavy key's R rid (in the
        // counted layout the histogram pass does); S tiles read the rids in
        // a later launch
        const bool rec = (I.heavy != 0) && (td.rel == 0) && (td.cap != 0) && heavy_keys_of(I) != 0;
        const uint32_t rn = min(ROUND, td.end - o);
        if ((heavy || rec) && heavy_inst != td.inst) {
            // the last lookups in the old table were before the previous
            // round's ranking barrier
            load_heavy(I, hk, kHeavyWords);
            heavy_inst = td.inst;
            __syncthreads();
        }
        uint32_t dg[IPT], rk[IPT];
        if (heavy) {
            // in phases, so that the LDS reads of all IPT tuples overlap:
            // table lookups (keyed by the SuperFastHash that also gives the
            // digit), then each slot's rid and presence word -- read for
            // every tuple, slot 0 for the others, so that no read waits
            // behind a branch -- then one rank atomic per wave for the heavy
            // bucket F, then the digit atomics below
            int hs[IPT];
            uint32_t hh[IPT];
#pragma unroll
            for (int j = 0; j < IPT; j++) {
                const uint32_t li = j * NT + threadIdx.x;
                hh[j] = sfh(v[j].x);
                hs[j] = heavy_find(hk, v[j].x, hh[j], li < rn && v[j].x != kEmptyKey);
                // lookups in groups: the empty asm (memory clobber) keeps
                // the next group's reads below this point, which bounds the
                // registers in flight (all 12 at once spill)
                if ((j + 1) % kHeavyLookupGroup == 0)
                    for (int u = j + 1 - kHeavyLookupGroup; u <= j; u++) asm volatile("" : "+v"(hs[u])::"memory");
            }
            uint32_t htot = 0;
#pragma unroll
            for (int j = 0; j < IPT; j++) {
                const uint32_t li = j * NT + threadIdx.x;
                const uint32_t sl = hs[j] >= 0 ? (uint32_t)hs[j] : 0u;
                const uint32_t pw = hr[kHeavySlots + (sl >> 5)], rid = hr[sl];
                // a heavy key with no R tuple: the tuple is dropped, as one
                // with the empty key 0 (it was not counted either); selects
                // only, so that no read is moved into a branch of its own
                const uint32_t present = __builtin_amdgcn_ubfe(pw, sl, 1u);   // bit sl & 31
                const uint32_t dig = (hh[j] >> shift) & mask;                // sfh's bits 0..18 are sfh_low's
                // (the selects on loaded values are written as arithmetic:
                // as ?: the compiler makes them branches, and moves the
                // presence read into the branch with a wait of its own)
                const uint32_t hd = DROP ^ ((0u - present) & (F ^ DROP));                    // present ? F : DROP
                const uint32_t found = (uint32_t)hs[j] >> 31 ^ 1u;                          // hs[j] >= 0
                const uint32_t inner = dig ^ ((0u - found) & (hd ^ dig));                   // found ? hd : dig
                const uint32_t d = inner | (0u - (uint32_t)!((li < rn) & (v[j].x != kEmptyKey)));   // DROP when not valid
                dg[j] = d;
                v[j].x = d == F ? rid : v[j].x;   // heavy: bucket F, staged as (R rid, S rid)
                htot += (uint32_t)__popcll(__ballot(d == F));
            }
            if (htot) {
                uint32_t hb = 0;
                if (lane_id() == 0) hb = atomicAdd(&lst[F], htot);
                hb = __shfl(hb, 0, kWave);
                // the ballots again (cheaper than holding twelve 64-bit
                // masks in registers across the atomic)
#pragma unroll
                for (int j = 0; j < IPT; j++) {
                    const uint64_t hm = __ballot(dg[j] == F);
                    if (dg[j] == F) rk[j] = hb + mbcnt(hm);
                    hb += (uint32_t)__popcll(hm);
                }
            }
        } else {
            if (rec) {
                int hs[IPT];
#pragma unroll
                for (int j = 0; j < IPT; j++) {
                    const uint32_t li = j * NT + threadIdx.x;
                    hs[j] = heavy_find(hk, v[j].x, sfh(v[j].x), li < rn && v[j].x != kEmptyKey);
                    if ((j + 1) % kHeavyLookupGroup == 0)
                        for (int u = j + 1 - kHeavyLookupGroup; u <= j; u++) asm volatile("" : "+v"(hs[u])::"memory");
                }
#pragma unroll
                for (int j = 0; j < IPT; j++)
                    if (hs[j] >= 0) heavy_record_rid(I, (uint32_t)hs[j], v[j].y);
            }
#pragma unroll
            for (int j = 0; j < IPT; j++) {
                const uint32_t li = j * NT + threadIdx.x;
                dg[j] = (li < rn && v[j].x != kEmptyKey) ? ((sfh_low(v[j].x) >> shift) & mask) : DROP;
            }
        }
#pragma unroll
        for (int j = 0; j < IPT; j++)
            if (dg[j] != DROP && dg[j] != F) rk[j] = atomicAdd(&lst[dg[j]], 1u);
        if (more) load_tile_round<IPT>(nv, insts, tdn, on);
        __syncthreads();   // ranks counted; the previous round's write-out is done
        {
            // the previous round's counters, for the round after this one
            // (its ranking comes after this round's last barrier)
            uint32_t i0 = threadIdx.x;   // opaque, as in the write-out
            asm volatile("" : "+v"(i0));
            for (uint32_t i = i0; i < FM + 2; i += NT) lold[i] = 0;
        }
        // reserve each digit's run at the group cursor; the returned bases
        // are only needed after the scan and staging, which hide the latency
        uint32_t *cur = counts + td.cnt_off;
        uint32_t hres = 0;   // heavy pairs: reserved from the back of the output
        if (heavy && threadIdx.x == 0) {
            const uint32_t hn = lst[F];
            if (hn) hres = atomicAdd(&out_hi[td.inst], hn);
        }
        // the tile queue's next value, like the reservations: issued here,
        // its result due at the wait after the staging (one round after the
        // tile switch that asked for it, before the next switch reads it)
        uint32_t qv = 0;
        if (qwant) qv = atomicAdd(q, 1u);
        // Two adjacent digits per atomic: their cursors form one 64-bit word
        // (group counters start at even offsets), and the low cursor never
        // carries into the high one (cursors stay below 2^32).  Thread t
        // takes digits 2t, 2t + 1, 2(NT + t), 2(NT + t) + 1 (slot u of cb/cc
        // is digit 2((u >> 1) NT + t) + (u & 1)).
        uint32_t cc[FMAX / NT];
        unsigned long long cp[FMAX / NT / 2];
#pragma unroll
        for (uint32_t h = 0; h < FMAX / NT / 2; h++) {
            const uint32_t d0 = 2 * (h * NT + threadIdx.x);
            cp[h] = 0;
            cc[2 * h] = d0 < F ? lst[d0] : 0u;
            cc[2 * h + 1] = d0 + 1 < F ? lst[d0 + 1] : 0u;   // F == 1: no second digit
        }
        if (F >= 2) {
#pragma unroll
            for (uint32_t h = 0; h < FMAX / NT / 2; h++) {
                const uint32_t d0 = 2 * (h * NT + threadIdx.x);
                if (d0 < F && (cc[2 * h] | cc[2 * h + 1]))
                    cp[h] = atomicAdd(reinterpret_cast<unsigned long long *>(cur + d0),
                                      (unsigned long long)cc[2 * h] | ((unsigned long long)cc[2 * h + 1] << 32));
            }
        } else if (threadIdx.x == 0 && cc[0]) {
            cp[0] = atomicAdd(&cur[0], cc[0]);
        }
        const uint32_t total = block_exclusive_scan(lst, F + 1, wtmp);   // lst[d] = start of digit d
        const uint32_t hstart = lst[F];                                  // == total without heavy tuples
#pragma unroll
        for (int j = 0; j < IPT; j++) {
            if (dg[j] != DROP) {
                const uint32_t p = lst[dg[j]] + rk[j];
                stage[p] = v[j];
            }
        }
        // One unconditional wait for all vector memory operations: the next
        // round's loads and the reservations.  The compiler then knows those
        // loads are done at the next round's start and does not wait there --
        // a wait that would also cover this round's stores, issued after them.
        __builtin_amdgcn_s_waitcnt(0);
        if (qwant) {   // thread 0; the atomic completed with the wait above
            misc[1] = qv;
            qwant = false;
        }
        __syncthreads();   // staging has read the starts; lst[d] becomes digit d's base - start
#pragma unroll
        for (uint32_t u = 0; u < FMAX / NT; u++) {
            const uint32_t d = 2 * ((u >> 1) * NT + threadIdx.x) + (u & 1u);
            if (d < F) {
                uint32_t b = (uint32_t)(cp[u >> 1] >> ((u & 1u) * 32));
                if (td.cap != 0 && cc[u] != 0 && b + cc[u] > td.cell0 + (d + 1) * td.cap) {
                    // the cell is full: the run goes to the sink and the
                    // join is redone with counted partitions
                    b = td.rel ? I.s_sink : I.r_sink;
                    atomicOr(&err[td.inst], kErrCellOverflow);
                }
                bs[d] = b - lst[d];
            }
        }
        if (heavy && threadIdx.x == 0) {
            // heavy pairs fill the output from its end downwards
            const uint32_t hn = total - hstart;
            misc[0] = hn ? I.nS - hres - hn : 0u;
        }
        __syncthreads();
        scatter_write_out(stage, bs, shift, mask, hstart, heavy ? total : hstart, td.rel ? I.Sp : I.Rp, I.out, misc[0], I.nS,
                          &err[td.inst]);
        if (!more) break;
        par = !par;
        if (tn != t) {   // next tile: the one after it was taken a round ago at least
            tnext = x + kScatterGroups * misc[1];
            qwant = threadIdx.x == 0;
        }
        t = tn;
        td = tdn;
        o = on;
    }
}

// ---------------------------------------------------------------------------
// Heavy-key detection (the skew step SPID handles with bank-set / rank-set
// replication, glb_chipwise_partition_count.c:251-255 and
// queryHostFunc.cc:1181-1419).  Sample blocks of S are counted exactly in an
// LDS table; keys seen at least kLocalMin times in a block are merged into a
// global candidate map, and k_heavy_select keeps candidates whose summed
// sample count reaches min_count.  The hot keys' R tuples are then held by
// every workgroup (the heavy table in LDS) and their S tuples are joined in
// whatever workgroup reads them: R replicated to all workgroups, S split by
// position -- the same multiset as the plain join.
// ---------------------------------------------------------------------------
constexpr uint32_t kSampleBlock = 8192;
constexpr uint32_t kGateMin = 8;          // pre-sample count of one key that opens the full sample
constexpr uint32_t kSampleSlots = 16384;
constexpr uint32_t kLocalMin = 4;
constexpr uint32_t kCandSlots = 16384;
static_assert(kCandSlots % 1024 == 0, "k_heavy_select reads the candidate map 1024 slots at a time");

constexpr uint32_t kAccSlots = 2048;      // a workgroup's candidates over its sample blocks (LDS)

// Add cnt sampled occurrences of key to the job's global candidate map.
// A slot holding a key never changes, so a plain read that sees one is
// exact; only a slot seen empty (possibly stale) takes the CAS.  Phase 0
// needs the sum for its gate; phase 1 adds without a return value, so
// nothing waits for it.
__device__ __forceinline__ void cand_merge(uint32_t *cand_keys, uint32_t *cand_cnt, uint32_t *found, uint32_t key,
                                           uint32_t cnt, uint32_t phase)
{
    uint32_t s = (key * 0x9E3779B1u) >> (32 - 14);
    for (uint32_t step = 0; step < kCandSlots; step++) {
        uint32_t old = gld32(cand_keys, s);
        if (old == kEmptyKey) old = atomicCAS(&cand_keys[s], kEmptyKey, key);
        if (old == kEmptyKey || old == key) {
            if (phase == 0) {
                // open the second phase only on a key seen often across the
                // pre-sample, not on one chance cluster in a single block
                if (atomicAdd(&cand_cnt[s], cnt) + cnt >= kGateMin) *found = 1u;
            } else {
                atomicAdd(&cand_cnt[s], cnt);
            }
            return;
        }
        s = (s + 1u) & (kCandSlots - 1u);
    }
}

__global__ __launch_bounds__(1024) void k_heavy_sample(const InstDesc *__restrict__ insts, const HeavyJob *__restrict__ jobs,
                                                      uint32_t njobs, uint32_t *__restrict__ cand_base, uint32_t phase,
                                                      uint32_t nvb)
{
    extern __shared__ __align__(16) uint32_t lds[];
    uint32_t *sk = lds;
    uint32_t *sc = lds + kSampleSlots;
    // The workgroup's candidates (keys seen kLocalMin times in a block) are
    // summed over its sample blocks of one job in LDS and merged into the
    // job's global map once: every block holds the same hot keys, and
    // merging each block's would put all workgroups' atomics on the same
    // few global words at once.
    uint32_t *ak = sc + kSampleSlots;
    uint32_t *ac = ak + kAccSlots;
    for (uint32_t i = threadIdx.x; i < 2 * kAccSlots; i += blockDim.x) ak[i] = 0;
    uint32_t acc_job = 0xFFFFFFFFu;   // job whose counts ak / ac hold (uniform)
    bool skip_job = false;            // phase 1: the job's pre-sample opened no gate
    // each thread flushes and clears its own accumulator slots
    auto flush = [&](uint32_t jb) {
        uint32_t *ck = cand_base + jobs[jb].cand_off, *cc = ck + kCandSlots, *fnd = cc + kCandSlots;
        for (uint32_t i = threadIdx.x; i < kAccSlots; i += blockDim.x) {
            if (ac[i]) cand_merge(ck, cc, fnd, ak[i], ac[i], phase);
            ak[i] = 0;
            ac[i] = 0;
        }
    };
    // workgroups loop over the phase's nvb sample blocks (a grid of at most
    // one workgroup per CU: when phase 1 has nothing to do, its blocks exit
    // without each being dispatched)
    for (uint32_t vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
        __syncthreads();   // the previous sample block's table reads are done
        // this block's job: last job whose first_block[phase] <= vb
        uint32_t lo = 0, hi = njobs;
        while (hi - lo > 1) {
            const uint32_t m = (lo + hi) >> 1;
            if (jobs[m].first_block[phase] <= vb) lo = m; else hi = m;
        }
        if (lo != acc_job) {   // a new job: the previous one's counts go to its map
            if (acc_job != 0xFFFFFFFFu) flush(acc_job);
            acc_job = lo;
            skip_job = phase == 1 && *(cand_base + jobs[lo].cand_off + 2 * kCandSlots) == 0u;
        }
        // no candidate in the pre-sample: no skew worth the rest
        if (skip_job) continue;
        // the job's fields are read where they lie (a local copy indexed by
        // phase would live in scratch memory)
        const HeavyJob &job = jobs[lo];
        const InstDesc &I = insts[job.inst];
        // sample block of this workgroup: phase 0 takes blocks 0, kPreSample,
        // 2 kPreSample, ...; phase 1 the others
        const uint32_t l = vb - job.first_block[phase];
        const uint32_t k = phase == 0 ? l * kPreSample : (l / (kPreSample - 1)) * kPreSample + l % (kPreSample - 1) + 1;
        if (k >= job.nblocks) continue;
        const uint2 *S = I.S;   // read once: reloaded per key, each key's load would wait for it
        const uint64_t b = (uint64_t)k * job.stride * kSampleBlock;
        const uint64_t e = min<uint64_t>(I.nS, b + kSampleBlock);
        // all of the block's keys are loaded first (kSampleBlock / 1024 per
        // thread), so the loads overlap instead of waiting one by one; the
        // table is cleared while they are in flight
        constexpr uint32_t KPT = kSampleBlock / 1024;
        uint32_t keys[KPT];
#pragma unroll
        for (uint32_t u = 0; u < KPT; u++) {
            const uint64_t i = b + u * 1024 + threadIdx.x;
            keys[u] = (i < e) ? gld(S, i).x : kEmptyKey;
        }
        for (uint32_t i = threadIdx.x; i < kSampleSlots; i += blockDim.x) { sk[i] = 0; sc[i] = 0; }
        __syncthreads();
#pragma unroll
        for (uint32_t u = 0; u < KPT; u++) {
            const uint32_t key = keys[u];
            if (key == kEmptyKey) continue;
            uint32_t s = (key * 0x9E3779B1u) >> (32 - 14);
            for (uint32_t step = 0; step < kSampleSlots; step++) {
                uint32_t old = atomicCAS(&sk[s], kEmptyKey, key);
                if (old == kEmptyKey || old == key) { atomicAdd(&sc[s], 1u); break; }
                s = (s + 1u) & (kSampleSlots - 1u);
            }
        }
        __syncthreads();
        uint32_t *cand_keys = cand_base + job.cand_off, *cand_cnt = cand_keys + kCandSlots;
        uint32_t *found = cand_cnt + kCandSlots;
        // the counts are read all at once (one wait), then the few keys seen
        // kLocalMin times go to the accumulator
        constexpr uint32_t SPT = kSampleSlots / 1024;
        uint32_t cnts[SPT];
#pragma unroll
        for (uint32_t u = 0; u < SPT; u++) cnts[u] = sc[u * 1024 + threadIdx.x];
#pragma unroll
        for (uint32_t u = 0; u < SPT; u++) {
            const uint32_t cnt = cnts[u];
            if (cnt < kLocalMin) continue;
            const uint32_t key = sk[u * 1024 + threadIdx.x];
            uint32_t s = (key * 0x9E3779B1u) >> (32 - 11);
            bool placed = false;
            for (uint32_t step = 0; step < kAccSlots; step++) {
                const uint32_t old = atomicCAS(&ak[s], kEmptyKey, key);
                if (old == kEmptyKey || old == key) {
                    atomicAdd(&ac[s], cnt);
                    placed = true;
                    break;
                }
                s = (s + 1u) & (kAccSlots - 1u);
            }
            if (!placed) cand_merge(cand_keys, cand_cnt, found, key, cnt, phase);   // accumulator full
        }
    }
    __syncthreads();   // the last block's accumulation is done
    if (acc_job != 0xFFFFFFFFu) flush(acc_job);
}

__device__ __forceinline__ void bitonic_sort_lds(uint64_t *a, uint32_t n /* power of two */)
{
    for (uint32_t k = 2; k <= n; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
                const uint32_t ixj = i ^ j;
                if (ixj > i) {
                    const uint64_t x = a[i], y = a[ixj];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) { a[i] = y; a[ixj] = x; }
                }
            }
            __syncthreads();
        }
    }
}

// Lane exchange with lane ^ M: DPP for masks 1 and 2 (quad permutes) and 8
// (rotating a 16-lane row by 8), a swizzle for the other masks below 32, a
// permute through the LDS crossbar for 32.
template <uint32_t M>
__device__ __forceinline__ uint32_t xlane(uint32_t v)
{
    if constexpr (M == 1) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);   // quad_perm 1,0,3,2
    else if constexpr (M == 2) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);   // quad_perm 2,3,0,1
    else if constexpr (M == 8) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false);  // row_ror:8
    else if constexpr (M < 32) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (M << 10));   // xor M in 32 lanes
    else return (uint32_t)__builtin_amdgcn_ds_bpermute((int)((lane_id() ^ M) << 2), (int)v);
}

template <uint32_t M, int E>
__device__ __forceinline__ void xlane_stage(uint64_t *x, uint32_t k, uint32_t lane, bool up)
{
    if (M * (uint32_t)E > (k >> 1)) return;   // no stage j = M E in this merge
    const bool keep_min = ((lane & M) == 0) == up;
#pragma unroll
    for (int r = 0; r < E; r++) {
        const uint64_t y = ((uint64_t)xlane<M>((uint32_t)(x[r] >> 32)) << 32) | xlane<M>((uint32_t)x[r]);
        x[r] = keep_min ? (x[r] < y ? x[r] : y) : (x[r] < y ? y : x[r]);
    }
}

// The same network on a[0, 1024 E) with the elements in registers: element
// e = t E + r is register r of thread t.  A stage whose partner distance j
// is below E runs inside each thread, below 64 E across the lanes of a wave
// (lane exchanges), and only the rest goes through LDS with a barrier: for
// 16K elements 10 barrier stages instead of 105.  Between those stages the
// elements sit in LDS at e + e / 16 (a pad word every 16): a thread's run
// of E then starts 17 words after its neighbour's instead of 16, which
// keeps the lanes of a wave off the same banks.  a must hold 17/16 of the
// elements; they are read and left in a[0, 1024 E) as given.
template <int E>
__device__ void bitonic_sort_regs(uint64_t *a)
{
    constexpr uint32_t N = 1024u * E;
    const uint32_t t = threadIdx.x, lane = lane_id();
    auto P = [](uint32_t e) { return e + (e >> 4); };
    uint64_t x[E];
#pragma unroll
    for (int r = 0; r < E; r++) x[r] = a[t * E + r];
    __syncthreads();   // every element is read before the padded layout overwrites them
    for (uint32_t k = 2; k <= N; k <<= 1) {
        uint32_t j = k >> 1;
        if (j >= 64u * E) {
#pragma unroll
            for (int r = 0; r < E; r++) a[P(t * E + r)] = x[r];
            __syncthreads();
            for (; j >= 64u * E; j >>= 1) {
                for (uint32_t p = t; p < N / 2; p += 1024u) {
                    const uint32_t lo = ((p & ~(j - 1u)) << 1) | (p & (j - 1u)), hi = lo + j;
                    const bool up = (lo & k) == 0;
                    const uint64_t u = a[P(lo)], v = a[P(hi)];
                    if ((u > v) == up) {
                        a[P(lo)] = v;
                        a[P(hi)] = u;
                    }
                }
                __syncthreads();
            }
            // each thread reads back only its own elements
#pragma unroll
            for (int r = 0; r < E; r++) x[r] = a[P(t * E + r)];
        }
        // partner lane ^ M for the stage j = M E (largest first), each only
        // when j <= k / 2
        const bool up = ((t * E) & k) == 0;   // k > j >= E: the same for every r
        xlane_stage<32, E>(x, k, lane, up);
        xlane_stage<16, E>(x, k, lane, up);
        xlane_stage<8, E>(x, k, lane, up);
        xlane_stage<4, E>(x, k, lane, up);
        xlane_stage<2, E>(x, k, lane, up);
        xlane_stage<1, E>(x, k, lane, up);
#pragma unroll
        for (int jj = E / 2; jj >= 1; jj >>= 1) {
            if ((uint32_t)jj < k) {
#pragma unroll
                for (int r = 0; r < E; r++) {
                    if ((r & jj) == 0) {
                        const bool upr = ((t * E + r) & k) == 0;
                        const uint64_t u = x[r], v = x[r | jj];
                        const bool sw = (u > v) == upr;
                        x[r] = sw ? v : u;
                        x[r | jj] = sw ? u : v;
                    }
                }
            }
        }
    }
    __syncthreads();   // every padded element is read back before the plain layout overwrites them
#pragma unroll
    for (int r = 0; r < E; r++) a[t * E + r] = x[r];
    __syncthreads();
}

__global__ __launch_bounds__(1024) void k_heavy_select(const InstDesc *__restrict__ insts, const HeavyJob *__restrict__ jobs,
                                                       const uint32_t *__restrict__ cand_base,
                                                       uint32_t *__restrict__ nheavy)
{
    constexpr uint32_t kSel = 4096;   // candidates sorted by count (more are rare; the rest is dropped)
    constexpr uint32_t kWaveKeys = 256;
    __shared__ uint32_t n, placed;
    __shared__ uint64_t ord[kSel];                 // count << 32 | key
    __shared__ __align__(16) uint32_t tab[kHeavySlots];
    __shared__ uint32_t bfill[kHeavySlots / 4];   // slots taken (or asked for) per bucket
    const HeavyJob job = jobs[blockIdx.x];
    const InstDesc &I = insts[job.inst];
    const uint32_t *cand_keys = cand_base + job.cand_off, *cand_cnt = cand_keys + kCandSlots;
    const uint32_t min_count = job.min_count;
    // the pre-sample opened no gate and no key's pre-sample count can reach
    // min_count: no heavy key (the table and count stay zero)
    if (cand_cnt[kCandSlots] == 0 && min_count >= kGateMin) return;
    if (threadIdx.x == 0) {
        n = 0;
        placed = 0;
    }
    for (uint32_t i = threadIdx.x; i < kHeavySlots; i += blockDim.x) tab[i] = kEmptyKey;
    for (uint32_t i = threadIdx.x; i < kHeavySlots / 4; i += blockDim.x) bfill[i] = 0;
    for (uint32_t i = threadIdx.x; i < kSel; i += blockDim.x) ord[i] = 0;
    __syncthreads();
    {
        // all of a thread's slots are loaded first (one wait, not one per
        // slot), then the candidates that reach min_count are appended
        constexpr uint32_t SPT = kCandSlots / 1024;
        uint32_t ck[SPT], cc[SPT];
#pragma unroll
        for (uint32_t u = 0; u < SPT; u++) {
            ck[u] = gld32(cand_keys, u * 1024 + threadIdx.x);
            cc[u] = gld32(cand_cnt, u * 1024 + threadIdx.x);
        }
#pragma unroll
        for (uint32_t u = 0; u < SPT; u++) {
            if (ck[u] == kEmptyKey || cc[u] < min_count) continue;
            const uint32_t k = atomicAdd(&n, 1u);
            if (k < kSel) ord[k] = ((uint64_t)cc[u] << 32) | ck[u];
        }
    }
    __syncthreads();
    const uint32_t nsel = min(n, kSel);
    if (nsel == 0) return;   // the table and count stay zero
    // hottest first: placement in rounds of kWaveKeys keys in descending
    // (count, key) order, so that a key left out for two full buckets is
    // one of the coldest selected (a hot key left ordinary can overflow a
    // cell of the partition layout).  Up to 1024 candidates (the usual
    // case) each thread ranks one by counting the larger ones, read as LDS
    // broadcasts, and stores it at its rank; more are sorted ascending by a
    // bitonic network (55 barrier stages already at 1024 elements).
    uint32_t m = 1;
    while (m < nsel) m <<= 1;
    const bool ranked = nsel <= 1024;
    if (ranked) {
        const uint64_t mine = threadIdx.x < nsel ? ord[threadIdx.x] : 0ull;
        uint32_t rk = 0;
        for (uint32_t j = 0; j < nsel; j++) rk += ord[j] > mine ? 1u : 0u;
        __syncthreads();   // every read of ord is done
        if (threadIdx.x < nsel) ord[rk] = mine;   // (count << 32 | key) is unique: the ranks are a permutation
        __syncthreads();
    } else {
        bitonic_sort_lds(ord, m);   // ord[nsel, m) are zeros: the keys end up in ord[m - nsel, m)
    }
    const uint32_t nkeys = min(nsel, kHeavyMaxKeys);
    // 2-choice placement, one key per thread per round: a slot of the
    // emptier of the key's two 4-slot buckets (as seen now), else of the
    // other; a key with both buckets full stays an ordinary key.  Keys are
    // unique, so each sits in at most one slot.
    for (uint32_t r0 = 0; r0 < nkeys; r0 += kWaveKeys) {
        if (threadIdx.x < kWaveKeys && r0 + threadIdx.x < nkeys) {
            const uint32_t r = r0 + threadIdx.x;
            const uint32_t key = (uint32_t)ord[ranked ? r : m - 1 - r];
            const uint32_t h = sfh(key), b1 = heavy_b1(h), b2 = heavy_b2(h);
            const uint32_t first = bfill[b2] < bfill[b1] ? b2 : b1, second = first == b1 ? b2 : b1;
            uint32_t slot = atomicAdd(&bfill[first], 1u), b = first;
            if (slot >= 4 && second != first) {
                slot = atomicAdd(&bfill[second], 1u);
                b = second;
            }
            if (slot < 4) {
                tab[4 * b + slot] = key;
                atomicAdd(&placed, 1u);
            }
        }
        __syncthreads();
    }
    // the rids and presence bits stay as k_prep zeroed them: no R tuple yet
    for (uint32_t i = threadIdx.x; i < kHeavySlots; i += blockDim.x) gst32(I.heavy_keys, i, tab[i]);
    if (threadIdx.x == 0) nheavy[job.inst] = placed;
}

// ---------------------------------------------------------------------------
// k_make_items: the device work queue.  For every partition with tuples on
// both sides, ceil(|S_p| / smax) items split its S range; all instances of a
// batch append to one dense list.  With a cell layout the partition's tuples
// of that relation are the fills of its kScatterGroups cells, which the
// partition's SMap entry strings together; a cell's fill is clamped to its
// capacity (an overflowed cell has been flagged, and the join is redone).
// spart[inst] and rpart[inst] (zeroed before) receive the number of
// partitioned S and R tuples.  Grid (instances, partition chunks of kItemsNT).
// ---------------------------------------------------------------------------
// Partition p's position map from its cells' fills (clamped to the
// capacity); returns the partition's tuple count.  The group descriptors
// and all fills are loaded before any is used, so the loads overlap.
__device__ __forceinline__ uint32_t cell_map(const CellGroup *__restrict__ cg, const uint32_t *__restrict__ counts,
                                             uint32_t p, SMap &m)
{
    uint32_t cap[kScatterGroups], cb[kScatterGroups], cur[kScatterGroups];
#pragma unroll
    for (uint32_t g = 0; g < kScatterGroups; g++) {
        cap[g] = cg[g].cap;
        cb[g] = cg[g].cell0 + p * cap[g];
    }
#pragma unroll
    for (uint32_t g = 0; g < kScatterGroups; g++) cur[g] = cap[g] ? counts[cg[g].cnt_off + p] : cb[g];
    uint32_t n = 0;
#pragma unroll
    for (uint32_t g = 0; g < kScatterGroups; g++) {
        const uint32_t fill = min(cur[g] - cb[g], cap[g]);
        m.V[g] = n;
        m.D[g] = cb[g] - n;
        n += fill;
    }
    return n;
}

__device__ __forceinline__ void identity_map(SMap &m)
{
#pragma unroll
    for (uint32_t g = 0; g < kScatterGroups; g++) {
        m.V[g] = g ? 0xFFFFFFFFu : 0u;
        m.D[g] = 0;
    }
}

// 256 partitions per block: the item words (160 bytes per partition) are
// written by 16 blocks per instance at 2^12 partitions, not 4.
constexpr uint32_t kItemsNT = 256;

__global__ __launch_bounds__(kItemsNT) void k_make_items(const InstDesc *__restrict__ insts, const uint32_t *__restrict__ pstart,
                                                     const CellGroup *__restrict__ cells,
                                                     const uint32_t *__restrict__ counts,
                                                     uint32_t smax, WorkItem *__restrict__ items,
                                                     uint32_t *__restrict__ nitems, uint32_t *__restrict__ spart,
                                                     uint32_t *__restrict__ rpart)
{
    __shared__ uint32_t wtmp[kItemsNT / kWave + 1];
    __shared__ uint32_t cnt[kItemsNT];
    __shared__ uint32_t base;
    const uint32_t inst = blockIdx.x;
    const InstDesc &I = insts[inst];
    const uint32_t F = 1u << I.bits;
    const uint32_t *pr = pstart + I.pstartR, *ps = pstart + I.pstartS;
    const uint32_t p0 = blockIdx.y * kItemsNT;
    if (p0 >= F) return;   // whole block
    {
        const uint32_t p = p0 + threadIdx.x;
        uint32_t nr = 0, ns = 0, c = 0;
        SMap rm, sm;
        identity_map(rm);
        identity_map(sm);
        if (p < F) {
            nr = I.r_cells ? cell_map(cells + (I.r_cells - 1), counts, p, rm) : pr[p + 1] - pr[p];
            ns = I.s_cells ? cell_map(cells + (I.s_cells - 1), counts, p, sm) : ps[p + 1] - ps[p];
            c = (nr && ns) ? (ns + smax - 1) / smax : 0u;
        }
        cnt[threadIdx.x] = c;
        __syncthreads();
        const uint32_t tot = block_exclusive_scan(cnt, kItemsNT, wtmp);
        if (threadIdx.x == 0) base = tot ? atomicAdd(nitems, tot) : 0u;
        __syncthreads();
        uint32_t o = base + cnt[threadIdx.x];
        for (uint32_t k = 0; k < c; k++) {
            WorkItem w;
            w.inst = inst;
            w.part = p;
            if (I.s_cells) {
                w.sb = k * smax;
                w.se = min(ns, w.sb + smax);
            } else {
                w.sb = ps[p] + k * smax;
                w.se = min(ps[p + 1], w.sb + smax);
            }
            if (I.r_cells) {
                w.rb = 0;
                w.re = nr;
            } else {
                w.rb = pr[p];
                w.re = pr[p + 1];
            }
            w.pad[0] = w.pad[1] = 0;
            w.smap = sm;
            w.rmap = rm;
            items[o + k] = w;
        }
        // partitioned tuples of the instance: one atomic per wave and relation
        const uint32_t ws = wave_incl_sum(ns), wr = wave_incl_sum(nr);
        if (lane_id() == kWave - 1) {
            if (ws) atomicAdd(&spart[inst], ws);
            if (wr) atomicAdd(&rpart[inst], wr);
        }
    }
}

// ---------------------------------------------------------------------------
// k_join: persistent workgroups pop items from the queue.  Per item the
// algorithm is chosen by partition size and the instance's setting:
//   hash (a2a_srj_build.c + a2a_srj_probe_inner_mt.c): R_p is counting-
//     sorted into an LDS table of up to 8192 buckets (the reference's table
//     has 4096 slots) by a multiply-shift hash of the key; a probe reads two
//     16-bit bucket ends and scans the bucket (first equal key wins, as the
//     reference's linear probe).  Up to kTableCap R tuples per workgroup.
//     A larger R_p is split into Q sub-tables by high hash bits;
//   nested loop (nested_loop_join.c) for tiny R_p;
//   sort-merge (mpsm_join_sort_probe.c): R_p and S chunks bitonic-sorted in
//     LDS, then a per-thread merge.
// Output: each item reserves |S range| pairs of its instance's output once
// (items grow from the front; heavy pairs from the back), waves append with
// one LDS atomic each, and an item that matched fewer tuples leaves a hole
// that is recorded for the compaction step -- exact counts, no padding.
// ---------------------------------------------------------------------------
constexpr uint32_t kJoinNT = 1024;
constexpr uint32_t kJoinIPT = 4;   // S tuples per lane per probe chunk (smaller chunks: less end-of-item wait)
constexpr uint32_t kNLMax = 32;             // nested loop when |R_p| <= 32 (auto)
constexpr uint32_t kBucketBits = 13;
constexpr uint32_t kBuckets = 1u << kBucketBits;   // at most 8192 buckets (16-bit ends: 16 KB)
constexpr uint32_t kTableCap = 17400;       // 8 B per tuple + 16 KB of bucket ends
constexpr uint32_t kSubFill = 16384;        // mean sub-table fill when R_p needs several tables
constexpr uint32_t kResplits = 2;           // sub-table overflows re-split (2Q) before sort-merge
constexpr uint32_t kRangeSortMax = 64;      // sort-merge: longest key-range bucket sorted in place
constexpr uint32_t kScan = 4;               // bucket entries read unconditionally per probe
constexpr uint32_t kBigList = 3584;         // sort-merge: key-range buckets of 3+ entries listed (u16, LDS)
// byte offset of that list: after the table ((kTableCap + kScan) pairs) and
// its bucket ends (7 + 1 + kBuckets halves, the ends 16-byte aligned)
constexpr size_t kBigListOff = ((2 * ((size_t)kTableCap + kScan) * 4 + (kBuckets + 8) * 2) + 15) / 16 * 16;
constexpr uint32_t kSortR = 16384;          // sort-merge: R run sorted in LDS (u64 words)
constexpr uint32_t kSortTail = 1024;        // sort-merge: a second, short run in the same pass
constexpr uint32_t kSortPad = kSortR + kSortR / 16;   // the register sort's padded layout
static_assert(kJoinNT == 1024, "bitonic_sort_regs maps 1024 threads");
constexpr uint32_t kNLChunk = 8192;

struct JoinShared {
    uint32_t item;   // the item to process next
    uint32_t item2;  // the one after it (the queue runs two items ahead)
    uint32_t chunk;  // the probe's next free chunk
    uint32_t fill;   // pairs written in the item's reservation
    uint32_t base;   // reservation start
    uint32_t res;    // reservation length
    uint32_t kmin;   // sort-merge: the smallest and largest R key of the partition
    uint32_t kmax;
    uint32_t big;    // sort-merge: some key-range bucket is too long to sort in place
    uint32_t nbig;   // sort-merge: key-range buckets of three or more entries listed
};

// Append one pair per lane whose `hit` is set.  Positions come from a
// ballot; one LDS atomic per wave instruction.
__device__ __forceinline__ void wave_emit(bool hit, uint32_t r_rid, uint32_t s_rid, JoinShared *sh, uint2 *out,
                                          uint32_t *err_i)
{
    const uint64_t m = __ballot(hit);
    if (m == 0) return;
    uint32_t wb = 0;
    if (lane_id() == 0) wb = atomicAdd(&sh->fill, (uint32_t)__popcll(m));
    wb = __shfl(wb, 0, kWave);
    if (hit) {
        const uint32_t pos = wb + mbcnt(m);
        if (pos < sh->res) gst(out, sh->base + pos, make_uint2(r_rid, s_rid));
        else atomicOr(err_i, kErrOutOverflow);
    }
}

// Append the hits of a whole round (bit j of hitm: pair p[j] of this lane
// is a hit): one LDS atomic per wave for all N probes.
template <int N>
__device__ __forceinline__ void wave_emit_round(uint32_t hitm, const uint2 *p, JoinShared *sh, uint2 *out,
                                                uint32_t *err_i)
{
    uint64_t m[N];
    uint32_t tot = 0;
#pragma unroll
    for (int j = 0; j < N; j++) {
        m[j] = __ballot((hitm >> j) & 1u);
        tot += (uint32_t)__popcll(m[j]);
    }
    if (tot == 0) return;
    uint32_t wb = 0;
    if (lane_id() == 0) wb = atomicAdd(&sh->fill, tot);
    wb = __shfl(wb, 0, kWave);
    const uint32_t res = sh->res, base = sh->base;
#pragma unroll
    for (int j = 0; j < N; j++) {
        if ((hitm >> j) & 1u) {
            const uint32_t pos = wb + mbcnt(m[j]);
            if (pos < res) gst(out, base + pos, p[j]);
            else atomicOr(err_i, kErrOutOverflow);
        }
        wb += (uint32_t)__popcll(m[j]);
    }
}

// A position map (SMap) in scalar registers.  Vector-typed values, so that
// the selection stays in registers (an array here ends up in scratch memory,
// indexed).
typedef uint32_t u32x8_t __attribute__((ext_vector_type(8)));
struct SPos {
    u32x8_t V, D;
    // Sp index of partition-relative S position v
    __device__ __forceinline__ uint32_t operator()(uint32_t v) const
    {
        uint32_t d = D[0];
        d = (v >= V[1]) ? D[1] : d;
        d = (v >= V[2]) ? D[2] : d;
        d = (v >= V[3]) ? D[3] : d;
        d = (v >= V[4]) ? D[4] : d;
        d = (v >= V[5]) ? D[5] : d;
        d = (v >= V[6]) ? D[6] : d;
        d = (v >= V[7]) ? D[7] : d;
        return v + d;
    }
};
static_assert(kScatterGroups == 8, "SPos holds 8 cells");

// Load one probe chunk of a wave: S positions cb + j * 64 + lane (j < IPT)
// below se.
template <int IPT>
__device__ __forceinline__ void load_chunk(uint2 *v, const uint2 *S, const SPos &sp, uint32_t cb, uint32_t se,
                                           uint32_t lane)
{
#pragma unroll
    for (int j = 0; j < IPT; j++) {
        const uint32_t si = cb + j * kWave + lane;
        v[j] = (si < se) ? gld(S, sp(si)) : make_uint2(kEmptyKey, 0u);
    }
}

// Load one probe round: S positions [s0, s0 + rn) of the item, IPT per
// thread (lane: the thread's index).
template <int IPT>
__device__ __forceinline__ void load_round(uint2 *v, const uint2 *S, const SPos &sp, uint32_t s0, uint32_t se,
                                           uint32_t lane)
{
#pragma unroll
    for (int j = 0; j < IPT; j++) {
        const uint32_t si = s0 + j * kJoinNT + lane;
        v[j] = (si < se) ? gld(S, sp(si)) : make_uint2(kEmptyKey, 0u);
    }
}

// An item's 40 words in scalar registers: through the scalar cache, or from
// a copy in LDS (uniform values, read by every lane).
struct ItemRegs {
    u32x8_t h;        // inst, part, sb, se, rb, re, -, -
    SPos s, r;
    __device__ __forceinline__ void load_global(const WorkItem *items, uint32_t it)
    {
        typedef const __attribute__((address_space(4))) u32x8_t const_u32x8;
        const_u32x8 *m = reinterpret_cast<const_u32x8 *>(reinterpret_cast<uintptr_t>(items + it));
        h = m[0];
        s.V = m[1];
        s.D = m[2];
        r.V = m[3];
        r.D = m[4];
    }
    static __device__ __forceinline__ u32x8_t lds8(const uint32_t *p)
    {
        const u32x8_t x = *reinterpret_cast<const u32x8_t *>(p);
        u32x8_t y;
#pragma unroll
        for (int k = 0; k < 8; k++) y[k] = __builtin_amdgcn_readfirstlane(x[k]);
        return y;
    }
    __device__ __forceinline__ void load_lds(const uint32_t *p)
    {
        h = lds8(p);
        s.V = lds8(p + 8);
        s.D = lds8(p + 16);
        r.V = lds8(p + 24);
        r.D = lds8(p + 32);
    }
};
static_assert(kItemWords == 40 && kItemWords <= kWave, "ItemRegs holds 40 words, one wave copies them");

// kRange: with the sort-merge of partitions that fit one table (key-range
// buckets).  The instantiation without it serves batches with no sort-merge
// instance; its registers are allocated for the hash and nested-loop paths
// alone (an alg 2 item there takes the run sort, so it stays correct).
template <bool kRange>
__global__ __launch_bounds__(kJoinNT) void k_join(const InstDesc *__restrict__ insts,
                                                  const WorkItem *__restrict__ items, uint32_t *__restrict__ ctrl,
                                                  uint32_t *__restrict__ out_lo, uint32_t *__restrict__ out_fill,
                                                  uint32_t *__restrict__ err, uint4 *__restrict__ holes,
                                                  uint32_t max_holes, uint4 *__restrict__ ext)
{
    extern __shared__ __align__(16) uint32_t lds[];
    __shared__ JoinShared sh;
    __shared__ __align__(16) uint32_t wtmp[kJoinNT / kWave + 1];
    constexpr uint32_t NT = kJoinNT, IPT = kJoinIPT, ROUND = NT * IPT;
    const uint32_t total_items = ctrl[0];
    // The queue runs two items ahead, so the next item's words can be
    // loaded during the current item (at its probe's start, with the first
    // round's tuples) and kept in LDS (nw, of item nw_item): the item start
    // then reads LDS instead of waiting for a miss to HBM.
    __shared__ __align__(32) uint32_t nw[kItemWords];
    __shared__ uint32_t nw_item;
    if (threadIdx.x == 0) {
        sh.item = atomicAdd(&ctrl[1], 1u);
        sh.item2 = atomicAdd(&ctrl[1], 1u);
        nw_item = 0xFFFFFFFFu;
    }
    for (;;) {
        __syncthreads();
        const uint32_t it = __builtin_amdgcn_readfirstlane(sh.item);
        if (it >= total_items) break;
        // The hash path takes the next item's index and this item's output
        // reservation by atomics issued right before an explicit wait that
        // it makes anyway (at its probe's start); no wait on their result
        // registers lingers to the next item's start, where it would also
        // cover (vector memory completes in order) this item's last pair
        // stores.  The other paths take them at their start and end.
        ItemRegs ir;
        if (__builtin_amdgcn_readfirstlane(nw_item) == it) ir.load_lds(nw);
        else ir.load_global(items, it);
        struct { uint32_t inst, sb, se, rb, re; } w = {ir.h[0], ir.h[2], ir.h[3], ir.h[4], ir.h[5]};
        const SPos &spos = ir.s, &rpos = ir.r;
        const InstDesc &I = insts[w.inst];
        const uint32_t rb = w.rb, re = w.re;
        const uint32_t nr = re - rb;
        // The algorithm by partition size: the nested loop for a tiny R
        // partition, else the hash join (in sub-tables when R_p is larger
        // than one LDS table); a partition the sub-tables cannot hold goes on
        // to sort-merge, which takes any size.
        uint32_t alg = I.alg;
        if (alg == 0 && nr <= kNLMax) alg = 3;
        uint32_t *erri = err + w.inst;
        // The table paths (hash, and sort-merge of a partition that fits one
        // table) take the next item from the queue and the item's output
        // reservation at their first table, the other paths at their start
        // (reservation) and end (queue).  `advanced` is set (uniformly) once
        // both are taken; exactly one reservation is made per item.
        const bool table_path = alg == 0 || (alg == 2 && nr <= kTableCap);
        bool advanced = false;
        bool run_sort = false;   // sort-merge by sorted runs, not key-range buckets
        if (threadIdx.x == 0) {
            sh.res = w.se - w.sb;
            sh.fill = 0;
            if (!table_path) sh.base = atomicAdd(&out_lo[w.inst], w.se - w.sb);
        }
        __syncthreads();

        for (;;) {   // one pass; two when the hash path hands the item to sort-merge
            if (alg == 3) {
                // ---------------- nested loop: R chunks in LDS; every S tuple
                // compares against every R key of the chunk (first match wins)
                uint32_t *rk = lds, *rv = lds + kNLChunk;
                for (uint32_t c0 = 0; c0 < nr; c0 += kNLChunk) {
                    const uint32_t cn = min(kNLChunk, nr - c0);
                    __syncthreads();
                    for (uint32_t i = threadIdx.x; i < cn; i += NT) { uint2 r = gld(I.Rp, rpos(rb + c0 + i)); rk[i] = r.x; rv[i] = r.y; }
                    __syncthreads();
                    for (uint32_t s0 = w.sb; s0 < w.se; s0 += ROUND) {
                        uint2 v[IPT];
                        load_round<IPT>(v, I.Sp, spos, s0, w.se, threadIdx.x);
#pragma unroll
                        for (int j = 0; j < (int)IPT; j++) {
                            bool hit = false;
                            uint32_t rr = 0;
                            if (v[j].x != kEmptyKey) {
                                for (uint32_t i = 0; i < cn; i++) {
                                    if (rk[i] == v[j].x) { rr = rv[i]; hit = true; break; }
                                }
                            }
                            wave_emit(hit, rr, v[j].y, &sh, I.out, erri);
                        }
                    }
                }
            } else if (alg == 2 && (!kRange || run_sort || nr > kTableCap)) {
                // ---------------- sort-merge in runs (mpsm_join_sort_probe.c),
                // for a partition larger than one table or with keys too
                // clustered for the key-range buckets below: the R
                // partition is bitonic-sorted by key in LDS (runs of kSortR);
                // each S tuple then takes the R entry at the lower bound of its
                // key, found by binary search -- the entry the reference's merge
                // of the two sorted runs (probe(), 246-272) pairs it with.  The
                // merge walk needs S sorted; the search does not, so S is probed
                // as it lies, IPT tuples per thread with their searches in
                // lockstep.  Adjacent equal keys in a sorted run are duplicate R
                // keys.
                // A partition of up to kSortR + kSortTail tuples is one pass: a
                // run of kSortR and a short second run (kSortR-tuple partitions
                // overflow by a little about half the time; a second pass would
                // read and search all of S again).
                uint64_t *ra = reinterpret_cast<uint64_t *>(lds), *ra2 = ra + kSortPad;
                for (uint32_t c0 = 0; c0 < nr;) {
                    const uint32_t cn = min(kSortR, nr - c0);
                    const uint32_t rest = nr - c0 - cn;
                    const uint32_t cn2 = rest <= kSortTail ? rest : 0u;
                    uint32_t np = 1, np2 = 1;
                    while (np < cn) np <<= 1;
                    while (np2 < cn2) np2 <<= 1;
                    __syncthreads();
                    for (uint32_t i = threadIdx.x; i < np; i += NT) {
                        if (i < cn) { uint2 r = gld(I.Rp, rpos(rb + c0 + i)); ra[i] = ((uint64_t)r.x << 32) | r.y; }
                        else ra[i] = ~0ull;
                    }
                    if (cn2) {
                        for (uint32_t i = threadIdx.x; i < np2; i += NT) {
                            if (i < cn2) { uint2 r = gld(I.Rp, rpos(rb + c0 + cn + i)); ra2[i] = ((uint64_t)r.x << 32) | r.y; }
                            else ra2[i] = ~0ull;
                        }
                    }
                    __syncthreads();
                    // each ends with a barrier
                    if (np == 16384) bitonic_sort_regs<16>(ra);
                    else if (np == 8192) bitonic_sort_regs<8>(ra);
                    else if (np == 4096) bitonic_sort_regs<4>(ra);
                    else bitonic_sort_lds(ra, np);
                    if (cn2) bitonic_sort_lds(ra2, np2);
                    // lower bound of key k in the sorted run x[0, n)
                    auto lower = [](const uint64_t *x, uint32_t n, uint32_t k) {
                        uint32_t lo = 0, hi = n;
                        while (lo < hi) {
                            const uint32_t m = (lo + hi) >> 1;
                            if ((uint32_t)(x[m] >> 32) < k) lo = m + 1; else hi = m;
                        }
                        return lo;
                    };
                    {
                        bool dup = false;
                        for (uint32_t i = threadIdx.x; i + 1 < cn; i += NT) {
                            const uint32_t k = (uint32_t)(ra[i] >> 32);
                            dup |= k != kEmptyKey && k == (uint32_t)(ra[i + 1] >> 32);
                        }
                        for (uint32_t i = threadIdx.x; i < cn2; i += NT) {   // the second run: itself and against the first
                            const uint32_t k = (uint32_t)(ra2[i] >> 32);
                            if (k == kEmptyKey) continue;
                            dup |= i + 1 < cn2 && k == (uint32_t)(ra2[i + 1] >> 32);
                            const uint32_t q = lower(ra, cn, k);
                            dup |= q < cn && (uint32_t)(ra[q] >> 32) == k;
                        }
                        // the tuples of the partition's earlier passes against
                        // this pass's runs: a key repeated across passes is a
                        // duplicate too (each pair of runs is compared once, when
                        // the later one is sorted)
                        for (uint32_t i = threadIdx.x; i < c0; i += NT) {
                            const uint32_t k = gld(I.Rp, rpos(rb + i)).x;
                            if (k == kEmptyKey) continue;
                            const uint32_t q = lower(ra, cn, k);
                            dup |= q < cn && (uint32_t)(ra[q] >> 32) == k;
                            if (cn2) {
                                const uint32_t q2 = lower(ra2, cn2, k);
                                dup |= q2 < cn2 && (uint32_t)(ra2[q2] >> 32) == k;
                            }
                        }
                        if (dup) atomicOr(erri, kErrDuplicateKey);
                    }
                    for (uint32_t s0 = w.sb; s0 < w.se; s0 += ROUND) {
                        uint2 v[IPT];
                        load_round<IPT>(v, I.Sp, spos, s0, w.se, threadIdx.x);
                        uint32_t lo[IPT], hi[IPT];
#pragma unroll
                        for (int j = 0; j < (int)IPT; j++) {
                            lo[j] = 0;
                            hi[j] = cn;
                        }
                        // floor(log2 cn) + 1 halvings close every [lo, hi)
                        for (uint32_t step = cn; step > 0; step >>= 1) {
#pragma unroll
                            for (int j = 0; j < (int)IPT; j++) {
                                if (lo[j] < hi[j]) {
                                    const uint32_t m = (lo[j] + hi[j]) >> 1;
                                    if ((uint32_t)(ra[m] >> 32) < v[j].x) lo[j] = m + 1;
                                    else hi[j] = m;
                                }
                            }
                        }
                        uint32_t hitm = 0;
                        uint2 pr[IPT];
#pragma unroll
                        for (int j = 0; j < (int)IPT; j++) {
                            pr[j] = v[j];
                            if (v[j].x != kEmptyKey && lo[j] < cn) {
                                const uint64_t e = ra[lo[j]];
                                if ((uint32_t)(e >> 32) == v[j].x) {
                                    hitm |= 1u << j;
                                    pr[j] = make_uint2((uint32_t)e, v[j].y);   // (R rid, S rid)
                                }
                            }
                        }
                        if (cn2) {   // keys not in the first run: the second
#pragma unroll
                            for (int j = 0; j < (int)IPT; j++) {
                                if ((hitm >> j) & 1u || v[j].x == kEmptyKey) continue;
                                const uint32_t q = lower(ra2, cn2, v[j].x);
                                if (q < cn2 && (uint32_t)(ra2[q] >> 32) == v[j].x) {
                                    hitm |= 1u << j;
                                    pr[j] = make_uint2((uint32_t)ra2[q], v[j].y);
                                }
                            }
                        }
                        wave_emit_round<IPT>(hitm, pr, &sh, I.out, erri);
                    }
                    c0 += cn + cn2;
                }
            } else {
                // ---------------- hash join (and sort-merge of a partition
                // that fits one table: RANGE, key-range buckets)
                bool to_smj = false;   // the item goes on to the run sort
                auto table_join = [&](auto range_tag) {
                    constexpr bool RANGE = decltype(range_tag)::value;

                    // LDS: (key, rid) pairs counting-sorted by bucket, then 16-bit
                    // bucket bounds: bend[0] = 0 and bend[b + 1] = end of bucket b, so
                    // bucket b is [bend[b], bend[b + 1]).  ends = bend + 1 starts on
                    // a 16-byte boundary, so that one thread's eight ends are one
                    // 128-bit word.
                    // One table of 8192 buckets (fast): the count pass counts into
                    // ends (16-bit, packed atomics) and the scan turns the counts
                    // into starts in place.  Otherwise the 32-bit counters live in
                    // the pair area during the count pass (the fill pass overwrites
                    // them) and the scan's starts are copied into ends.
                    uint2 *tkv = reinterpret_cast<uint2 *>(lds);   // kTableCap + kScan slots
                    uint32_t *bcnt = lds;                           // nb counters (count pass only)
                    uint16_t *bend = reinterpret_cast<uint16_t *>(lds + 2 * (kTableCap + kScan)) + 7;   // kBuckets + 1
                    uint16_t *ends = bend + 1;
                    uint32_t *ends_w = reinterpret_cast<uint32_t *>(ends);   // two ends a word
                    // A partition larger than one table is split by high hash bits
                    // into Q sub-tables of kSubFill tuples on average (the slack up
                    // to kTableCap covers the spread of their sizes).  One that still
                    // overflows splits the partition again, into 2Q sub-tables, from
                    // the start; after kResplits such splits it goes to sort-merge.
                    uint32_t Q = nr <= kTableCap ? 1u : (nr + kSubFill - 1) / kSubFill;
                    uint32_t splits = 0;
                    const uint32_t lb = min(I.table_log2, kBucketBits), nb = 1u << lb;   // buckets
                    const bool fast = nb == 8 * NT && nr <= kTableCap;   // (then Q == 1)
                    constexpr uint32_t RPT = (kTableCap + NT - 1) / NT;
                    // the bucket of a key: multiply-shift hashing, or (sort-merge)
                    // its place in the partition's key range [kmin, kmax] cut into
                    // nb equal ranges -- monotone in the key, so the buckets in
                    // order, each sorted, are the partition sorted by key
                    uint32_t rmin = 0, rscale = 0;
                    auto bkt = [&](uint32_t k) -> uint32_t {
                        if constexpr (RANGE) {
                            const uint32_t g = (uint32_t)(((uint64_t)(k - rmin) * rscale) >> 32);
                            return g < nb ? g : nb - 1u;   // a key outside the range: any bucket (it matches no entry)
                        } else {
                            return bucket_of(k, lb);
                        }
                    };
                    for (uint32_t q = 0; q < Q;) {
                        __syncthreads();
                        {
                            uint32_t i0 = threadIdx.x;   // opaque: see the bucket ends below
                            asm volatile("" : "+v"(i0));
                            if (fast) reinterpret_cast<uint4 *>(ends)[i0] = make_uint4(0u, 0u, 0u, 0u);
                            else
                                for (uint32_t i = i0; i < nb; i += NT) bcnt[i] = 0;
                        }
                        if (RANGE && threadIdx.x == 0) {
                            sh.kmin = 0xFFFFFFFFu;
                            sh.kmax = 0u;
                            sh.big = 0u;
                            sh.nbig = 0u;
                        }
                        __syncthreads();
                        // R_p fits one table (Q == 1): hold it in registers (<= RPT
                        // tuples per thread) and read it once; otherwise each
                        // sub-table reads R_p twice, keeping its own keys
                        uint2 r[RPT];
#pragma unroll
                        for (uint32_t u = 0; u < RPT; u++) r[u] = make_uint2(kEmptyKey, 0u);   // defined on every path
                        // One table of 8192 buckets: the queue and reservation
                        // atomics and the load of the next item's words are issued
                        // before R's loads, so that the count's wait for R (vector
                        // memory completes in issue order) covers them, instead of a
                        // wait of their own at the probe's start.
                        const bool early = fast && !advanced;
                        uint32_t e_nx = 0, e_rb = 0, e_pre = 0, e_it2 = total_items;
                        if (Q == 1) {
                            // opaque lane: the positions are formed here, not hoisted
                            // to the item's start (where one spills, and its reload
                            // waits, in order, for the previous item's pair stores)
                            uint32_t l0 = threadIdx.x;
                            asm volatile("" : "+v"(l0));
                            if (early) {
                                if (threadIdx.x < kItemWords) e_it2 = sh.item2;
                                if (e_it2 < total_items) e_pre = gld32(reinterpret_cast<const uint32_t *>(items + e_it2), l0);
                                if (threadIdx.x == 0) {
                                    e_nx = atomicAdd(&ctrl[1], 1u);
                                    e_rb = atomicAdd(&out_lo[w.inst], w.se - w.sb);
                                }
                            }
#pragma unroll
                            for (uint32_t u = 0; u < RPT; u++) {
                                const uint32_t i = rb + u * NT + l0;
                                if (i < re) r[u] = gld(I.Rp, rpos(i));
                            }
                            if constexpr (RANGE) {
                                // the key range: per-thread, per-wave, then one LDS
                                // atomic per wave (sh.kmin / kmax were set at the clear)
                                uint32_t lo = 0xFFFFFFFFu, hi = 0u;
#pragma unroll
                                for (uint32_t u = 0; u < RPT; u++)
                                    if (r[u].x != kEmptyKey) {
                                        lo = min(lo, r[u].x);
                                        hi = max(hi, r[u].x);
                                    }
#pragma unroll
                                for (uint32_t off = kWave / 2; off > 0; off >>= 1) {
                                    lo = min(lo, (uint32_t)__shfl_xor((int)lo, (int)off, kWave));
                                    hi = max(hi, (uint32_t)__shfl_xor((int)hi, (int)off, kWave));
                                }
                                if (lane_id() == 0) {
                                    atomicMin(&sh.kmin, lo);
                                    atomicMax(&sh.kmax, hi);
                                }
                                __syncthreads();
                                rmin = sh.kmin;
                                const uint32_t rmax = sh.kmax;
                                // scale = floor(nb 2^32 / (range)), so a key's bucket
                                // (k - kmin) scale >> 32 is below nb inside the range
                                rscale = rmin <= rmax ? (uint32_t)min<uint64_t>(((uint64_t)nb << 32) / ((uint64_t)(rmax - rmin) + 1u),
                                                                                0xFFFFFFFFull)
                                                      : 0u;
                            }
                            if (fast) {
                                // a packed 16-bit counter never carries: counts <= nr < 2^16
#pragma unroll
                                for (uint32_t u = 0; u < RPT; u++) {
                                    if (r[u].x == kEmptyKey) continue;
                                    const uint32_t b = bkt(r[u].x);
                                    atomicAdd(&ends_w[b >> 1], 1u << ((b & 1u) * 16u));
                                }
                            } else {
#pragma unroll
                                for (uint32_t u = 0; u < RPT; u++)
                                    if (r[u].x != kEmptyKey) atomicAdd(&bcnt[bkt(r[u].x)], 1u);
                            }
                            if (early) {
                                // the next item's words and index, this item's
                                // reservation (read after the barriers below)
                                if (e_it2 < total_items) nw[threadIdx.x] = e_pre;
                                if (threadIdx.x == 0) {
                                    if (e_it2 < total_items) nw_item = e_it2;
                                    sh.item = sh.item2;
                                    sh.item2 = e_nx;
                                    sh.base = e_rb;
                                }
                                advanced = true;
                            }
                        } else {
                            for (uint32_t i0 = rb; i0 < re; i0 += 4 * NT) {
                                uint2 c[4];
#pragma unroll
                                for (int u = 0; u < 4; u++) {
                                    const uint32_t i = i0 + u * NT + threadIdx.x;
                                    c[u] = (i < re) ? gld(I.Rp, rpos(i)) : make_uint2(kEmptyKey, 0u);
                                }
#pragma unroll
                                for (int u = 0; u < 4; u++) {
                                    if (c[u].x == kEmptyKey || sub_of(sfh(c[u].x), I.shift + I.bits, Q) != q) continue;
                                    atomicAdd(&bcnt[bucket_of(c[u].x, lb)], 1u);
                                }
                            }
                        }
                        __syncthreads();
                        uint32_t ntab;
                        if (fast) {
                            // thread t: buckets 8t .. 8t + 7, one 128-bit word of
                            // counts, scanned in registers; the wave prefix by DPP,
                            // the waves' totals through wtmp (one barrier)
                            uint32_t t0 = threadIdx.x;   // opaque: addresses formed here
                            asm volatile("" : "+v"(t0));
                            uint4 *ev = reinterpret_cast<uint4 *>(ends) + t0;
                            const uint4 c = *ev;
                            const uint32_t e[8] = {c.x & 0xFFFFu, c.x >> 16, c.y & 0xFFFFu, c.y >> 16,
                                                   c.z & 0xFFFFu, c.z >> 16, c.w & 0xFFFFu, c.w >> 16};
                            uint32_t sum = 0;
#pragma unroll
                            for (int k = 0; k < 8; k++) sum += e[k];
                            const uint32_t x = wave_incl_sum(sum);
                            const uint32_t wid = __builtin_amdgcn_readfirstlane(threadIdx.x) / kWave;
                            if (lane_id() == kWave - 1) wtmp[wid] = x;
                            __syncthreads();
                            uint32_t pre = 0, tot = 0;
#pragma unroll
                            for (uint32_t k = 0; k < NT / kWave; k++) {
                                const uint32_t v = wtmp[k];
                                tot += v;
                                pre += k < wid ? v : 0u;
                            }
                            uint32_t run = pre + x - sum, o[4];
#pragma unroll
                            for (int k = 0; k < 4; k++) {
                                o[k] = run | ((run + e[2 * k]) << 16);
                                run += e[2 * k] + e[2 * k + 1];
                            }
                            *ev = make_uint4(o[0], o[1], o[2], o[3]);   // ends[b] = start of bucket b
                            ntab = __builtin_amdgcn_readfirstlane(tot);
                        } else {
                            ntab = block_exclusive_scan(bcnt, nb, wtmp);   // bcnt[b] = start of bucket b
                        }
                        if (ntab > kTableCap) {   // uniform: ntab is the same in every thread
                            if (!advanced) {   // the queue and reservation atomics, as below
                                uint32_t nx = 0, rbx = 0;
                                if (threadIdx.x == 0) {
                                    nx = atomicAdd(&ctrl[1], 1u);
                                    rbx = atomicAdd(&out_lo[w.inst], w.se - w.sb);
                                }
                                __builtin_amdgcn_s_waitcnt(0);
                                if (threadIdx.x == 0) {
                                    sh.item = sh.item2;
                                    sh.item2 = nx;
                                    sh.base = rbx;
                                }
                                advanced = true;
                            }
                            // the earlier sub-tables' pairs are written again, from
                            // the reservation's start (every barrier since their
                            // probe has passed, so no wave still appends)
                            if (threadIdx.x == 0) sh.fill = 0;
                            if (splits == kResplits) {
                                to_smj = true;
                                break;
                            }
                            splits++;
                            Q *= 2;
                            q = 0;
                            continue;
                        }
                        // 16-bit fill cursors (starts < 2^16: ntab <= kTableCap); the
                        // cursor of bucket b is ends[b] = bend[b + 1]
                        if (!fast) {
                            // the lane index is made opaque so that the addresses are
                            // formed here: hoisted out of the item loop they spill,
                            // and the reload waits (in order) for the previous item's
                            // pair stores
                            uint32_t i0 = threadIdx.x;
                            asm volatile("" : "+v"(i0));
                            for (uint32_t i = i0; i < nb; i += NT) ends[i] = (uint16_t)bcnt[i];
                        }
                        if (threadIdx.x == 0) bend[0] = 0;
                        __syncthreads();
                        // fill pass: a packed atomic advances one 16-bit cursor of a
                        // word; it never carries, every cursor staying <= ntab
                        if (Q == 1) {
#pragma unroll
                            for (uint32_t u = 0; u < RPT; u++) {
                                if (r[u].x == kEmptyKey) continue;
                                const uint32_t b = bkt(r[u].x), sh16 = (b & 1u) * 16u;
                                const uint32_t pos = (atomicAdd(&ends_w[b >> 1], 1u << sh16) >> sh16) & 0xFFFFu;
                                tkv[pos] = r[u];
                            }
                        } else {
                            for (uint32_t i0 = rb; i0 < re; i0 += 4 * NT) {
                                uint2 c[4];
#pragma unroll
                                for (int u = 0; u < 4; u++) {
                                    const uint32_t i = i0 + u * NT + threadIdx.x;
                                    c[u] = (i < re) ? gld(I.Rp, rpos(i)) : make_uint2(kEmptyKey, 0u);
                                }
#pragma unroll
                                for (int u = 0; u < 4; u++) {
                                    if (c[u].x == kEmptyKey || sub_of(sfh(c[u].x), I.shift + I.bits, Q) != q) continue;
                                    const uint32_t b = bucket_of(c[u].x, lb), sh16 = (b & 1u) * 16u;
                                    const uint32_t pos = (atomicAdd(&ends_w[b >> 1], 1u << sh16) >> sh16) & 0xFFFFu;
                                    tkv[pos] = c[u];
                                }
                            }
                        }
                        if (threadIdx.x == 0) sh.chunk = NT / kWave;   // the probe's first chunks go to waves 0..15
                        __syncthreads();   // bend[b + 1] = end of bucket b
                        constexpr uint32_t CH = kWave * IPT;
                        const uint32_t nch = (w.se - w.sb + CH - 1) / CH;
                        const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x) / kWave;
                        const uint32_t ln = lane_id();
                        // (issued before the duplicate check, which hides their latency)
                        uint2 nv[IPT];
                        {
                            // the lane index is made opaque so that the first chunk's
                            // positions are mapped here, not hoisted to the item's
                            // start (where they would spill, and their reload wait,
                            // in order, behind every store in flight)
                            uint32_t lane = ln;
                            asm volatile("" : "+v"(lane));
                            load_chunk<IPT>(nv, I.Sp, spos, w.sb + wv * CH, w.se, lane);
                        }
                        if constexpr (RANGE) {
                            // sort each key-range bucket in place (its entries are
                            // contiguous): four or fewer in registers by a network,
                            // up to kRangeSortMax by insertion; a longer one (keys
                            // clustered in the range) sends the item to the run
                            // sort below.  Equal keys share a bucket and end up
                            // adjacent: duplicates (mpsm_join_sort_probe.c sorts each
                            // range bucket, 152-236; a2a_srj_build.c:265-271).
                            bool dup = false, big = false;
                            // one bucket of n >= 3 entries from st: eight or fewer by
                            // a network on (key, rid) as 64-bit words (slots past n
                            // sort last; Batcher's odd-even merge, 19 comparators),
                            // up to kRangeSortMax by insertion, longer ones send the
                            // item to the run sort.  Slots past the bucket may lie
                            // past the table's slack, in the bucket ends: read, never
                            // used.
                            auto sort_bucket = [&](uint32_t st, uint32_t n) {
                                if (n <= 8) {
                                    uint64_t e[8];
#pragma unroll
                                    for (uint32_t u = 0; u < 8; u++) {
                                        const uint2 t = tkv[st + u];
                                        e[u] = u < n ? ((uint64_t)t.x << 32) | t.y : ~0ull;
                                    }
                                    auto cx = [&](int a, int c) {
                                        const uint64_t x = e[a], y = e[c];
                                        e[a] = x < y ? x : y;
                                        e[c] = x < y ? y : x;
                                    };
                                    cx(0, 1); cx(2, 3); cx(4, 5); cx(6, 7);
                                    cx(0, 2); cx(1, 3); cx(4, 6); cx(5, 7);
                                    cx(1, 2); cx(5, 6);
                                    cx(0, 4); cx(1, 5); cx(2, 6); cx(3, 7);
                                    cx(2, 4); cx(3, 5);
                                    cx(1, 2); cx(3, 4); cx(5, 6);
#pragma unroll
                                    for (uint32_t u = 0; u < 8; u++)
                                        if (u < n) tkv[st + u] = make_uint2((uint32_t)(e[u] >> 32), (uint32_t)e[u]);
#pragma unroll
                                    for (uint32_t u = 1; u < 8; u++) dup |= u < n && (e[u] >> 32) == (e[u - 1] >> 32);
                                } else if (n <= kRangeSortMax) {
                                    for (uint32_t a = st + 1; a < st + n; a++) {   // insertion sort (rare)
                                        const uint2 t = tkv[a];
                                        uint32_t c = a;
                                        while (c > st && tkv[c - 1].x > t.x) {
                                            tkv[c] = tkv[c - 1];
                                            c--;
                                        }
                                        tkv[c] = t;
                                        dup |= c > st && tkv[c - 1].x == t.x;
                                    }
                                } else {
                                    big = true;
                                }
                            };
                            // Pass 1: thread t takes buckets 8t .. 8t + 7 (their ends
                            // one 128-bit word, the first start the word before's
                            // last end).  A bucket of two entries is sorted by one
                            // comparison; the longer ones are listed (one LDS atomic
                            // per wave).  Pass 2 deals the listed buckets to the
                            // threads evenly: with about two entries per bucket, a
                            // third of them need the network, and every lane ran it
                            // for each of its buckets before (a wave waits for its
                            // longest bucket).  If the list overflows (more than
                            // kBigList such buckets: keys in clusters), pass 3 sorts
                            // every thread's own long buckets again -- a sorted
                            // bucket sorts to itself, with the same duplicates.
                            uint16_t *blist = reinterpret_cast<uint16_t *>(reinterpret_cast<char *>(lds) + kBigListOff);
                            uint32_t b0 = threadIdx.x;   // opaque: addresses formed here
                            asm volatile("" : "+v"(b0));
                            {
                                uint32_t be[9];   // bucket 8t + k is [be[k], be[k + 1])
#pragma unroll
                                for (uint32_t k = 0; k < 9; k++) be[k] = 0;
                                if (b0 < (nb >> 3)) {   // nb >= 64, a multiple of 8
                                    const uint4 c = reinterpret_cast<const uint4 *>(ends)[b0];
                                    be[0] = bend[8 * b0];
                                    be[1] = c.x & 0xFFFFu;
                                    be[2] = c.x >> 16;
                                    be[3] = c.y & 0xFFFFu;
                                    be[4] = c.y >> 16;
                                    be[5] = c.z & 0xFFFFu;
                                    be[6] = c.z >> 16;
                                    be[7] = c.w & 0xFFFFu;
                                    be[8] = c.w >> 16;
                                }
                                // the pairs' entries read four buckets at a time, one wait each
#pragma unroll
                                for (uint32_t h = 0; h < 8; h += 4) {
                                    uint2 xa[4], ya[4];
#pragma unroll
                                    for (uint32_t k = 0; k < 4; k++) {
                                        const uint32_t st = be[h + k];
                                        if (be[h + k + 1] - st == 2) {
                                            xa[k] = tkv[st];
                                            ya[k] = tkv[st + 1];
                                        }
                                    }
#pragma unroll
                                    for (uint32_t k = 0; k < 4; k++) {
                                        const uint32_t st = be[h + k];
                                        if (be[h + k + 1] - st != 2) continue;
                                        const uint2 x = xa[k], y = ya[k];
                                        if (x.x > y.x || (x.x == y.x && x.y > y.y)) {
                                            tkv[st] = y;
                                            tkv[st + 1] = x;
                                        }
                                        dup |= x.x == y.x;
                                    }
                                }
                                uint32_t c3 = 0;
#pragma unroll
                                for (uint32_t k = 0; k < 8; k++) c3 += be[k + 1] - be[k] >= 3 ? 1u : 0u;
                                const uint32_t incl = wave_incl_sum(c3);
                                const uint32_t wtot = __builtin_amdgcn_readlane(incl, kWave - 1);
                                if (wtot) {   // uniform
                                    uint32_t wb = 0;
                                    if (lane_id() == kWave - 1) wb = atomicAdd(&sh.nbig, wtot);
                                    uint32_t pos = __builtin_amdgcn_readlane(wb, kWave - 1) + incl - c3;
#pragma unroll
                                    for (uint32_t k = 0; k < 8; k++) {
                                        if (be[k + 1] - be[k] < 3) continue;
                                        if (pos < kBigList) blist[pos] = (uint16_t)(8 * b0 + k);
                                        pos++;
                                    }
                                }
                            }
                            __syncthreads();   // the list is complete
                            const uint32_t nbig = sh.nbig;
                            for (uint32_t i = b0; i < min(nbig, kBigList); i += NT) {
                                const uint32_t bb = blist[i];
                                const uint32_t st = bend[bb];
                                sort_bucket(st, (uint32_t)bend[bb + 1] - st);
                            }
                            if (nbig > kBigList) {   // uniform
                                __syncthreads();   // pass 2 is done with every bucket
                                if (b0 < (nb >> 3))
                                    for (uint32_t bb = 8 * b0; bb < 8 * b0 + 8; bb++) {
                                        const uint32_t st = bend[bb], n = bend[bb + 1] - st;
                                        if (n >= 3) sort_bucket(st, n);
                                    }
                            }
                            if (dup) atomicOr(erri, kErrDuplicateKey);
                            if (big) sh.big = 1u;
                            __syncthreads();   // the buckets are sorted; sh.big is final
                            if (sh.big) {
                                // too long a bucket: the run sort takes the item
                                // (once the queue and reservation atomics are done,
                                // as an overflowing hash table does)
                                if (!advanced) {
                                    uint32_t nx = 0, rbx = 0;
                                    if (threadIdx.x == 0) {
                                        nx = atomicAdd(&ctrl[1], 1u);
                                        rbx = atomicAdd(&out_lo[w.inst], w.se - w.sb);
                                    }
                                    __builtin_amdgcn_s_waitcnt(0);
                                    if (threadIdx.x == 0) {
                                        sh.item = sh.item2;
                                        sh.item2 = nx;
                                        sh.base = rbx;
                                    }
                                    advanced = true;
                                }
                                to_smj = true;
                                break;
                            }
                        } else {
                            // duplicate R keys (a2a_srj_build.c:265-271): each entry against
                            // the later entries of its bucket, in phases (keys, bucket ends,
                            // next entry) so that the LDS reads of all entries overlap;
                            // buckets of more than two entries finish in a short loop
                            {
                                bool dup = false;
                                uint32_t kk[RPT], ee[RPT];
#pragma unroll
                                for (uint32_t u = 0; u < RPT; u++) {
                                    const uint32_t i = u * NT + threadIdx.x;
                                    kk[u] = (i < ntab) ? tkv[i].x : kEmptyKey;
                                }
#pragma unroll
                                for (uint32_t u = 0; u < RPT; u++) ee[u] = (kk[u] != kEmptyKey) ? bend[bucket_of(kk[u], lb) + 1] : 0u;
                                // the next kScan entries are read unconditionally (the
                                // table has kScan slack slots past its last entry), so
                                // that the reads overlap; only buckets longer than that
                                // finish in a loop
#pragma unroll
                                for (uint32_t u = 0; u < RPT; u++) {
                                    const uint32_t i = u * NT + threadIdx.x;
                                    if (i >= ntab) continue;
#pragma unroll
                                    for (uint32_t d = 1; d <= kScan; d++) {
                                        const uint32_t kd = tkv[i + d].x;   // i + d < ntab + kScan: inside the table
                                        dup |= (i + d < ee[u]) & (kd == kk[u]);
                                    }
                                }
#pragma unroll
                                for (uint32_t u = 0; u < RPT; u++) {
                                    const uint32_t i = u * NT + threadIdx.x;
                                    for (uint32_t j = i + kScan + 1; j < ee[u]; j++) dup |= (tkv[j].x == kk[u]);
                                }
                                if (dup) atomicOr(erri, kErrDuplicateKey);
                            }
                        }
                        // probe: the S range is cut into chunks of a wave's 64 x IPT
                        // tuples; wave i starts on chunk i and then takes chunks from
                        // a counter, so that waves that finish early take over the
                        // rest instead of waiting at the item's end for slow ones.
                        // The next chunk's loads are in flight while one is probed.
                        uint32_t cg = 0;   // lane 0: the chunk after this wave's first (none without a first)
                        if (ln == 0 && wv < nch) cg = atomicAdd(&sh.chunk, 1u);
                        uint32_t nx = 0, rbx = 0, pre = 0;
                        // the next item's words: wave 0, lane i loads word i
                        const uint32_t it2 = !advanced && threadIdx.x < kItemWords ? sh.item2 : total_items;
                        if (it2 < total_items) {
                            // the lane offset is made opaque so that the address is
                            // formed here, not hoisted out of the item loop (a
                            // 64-bit per-lane pointer there would spill)
                            uint32_t lane = threadIdx.x;
                            asm volatile("" : "+v"(lane));
                            pre = gld32(reinterpret_cast<const uint32_t *>(items + it2), lane);
                        }
                        if (!advanced && threadIdx.x == 0) {
                            nx = atomicAdd(&ctrl[1], 1u);
                            rbx = atomicAdd(&out_lo[w.inst], w.se - w.sb);
                        }
                        __builtin_amdgcn_s_waitcnt(0);   // enter with no loads in flight (see the wait below)
                        if (!advanced) {
                            if (it2 < total_items) nw[threadIdx.x] = pre;
                            if (threadIdx.x == 0) {
                                if (it2 < total_items) nw_item = it2;
                                sh.item = sh.item2;   // read at the next item's start
                                sh.item2 = nx;
                                sh.base = rbx;        // read by the pair emission
                            }
                            __syncthreads();
                            advanced = true;
                        }
                        uint32_t cn = wv < nch ? __builtin_amdgcn_readfirstlane(cg) : nch;   // this wave's next chunk
                        for (uint32_t c = wv; c < nch;) {
                            uint2 v[IPT];
#pragma unroll
                            for (int j = 0; j < (int)IPT; j++) v[j] = nv[j];
                            if (cn < nch) load_chunk<IPT>(nv, I.Sp, spos, w.sb + cn * CH, w.se, ln);
                            cg = 0;
                            if (ln == 0 && cn < nch) cg = atomicAdd(&sh.chunk, 1u);   // the chunk after that
                            uint32_t hitm = 0;
#pragma unroll
                            for (int j = 0; j < (int)IPT; j++) {
                                // branch-free: key 0 and padding tuples simply find
                                // no entry (the table holds no key 0), nor does a key
                                // of another sub-table of a split partition; table
                                // keys are unique (duplicates are an error), so at
                                // most one entry matches
                                const uint32_t k = v[j].x;
                                const uint32_t bk = bkt(k);
                                const uint32_t st = bend[bk], n = bend[bk + 1] - st;
                                bool hit = false;
                                uint32_t rr = 0;
#pragma unroll
                                for (uint32_t u = 0; u < kScan; u++) {
                                    const uint2 e = tkv[st + u];   // the table has kScan padding slots
                                    const bool m = (u < n) && (e.x == k);
                                    rr = m ? e.y : rr;
                                    hit = hit || m;
                                }
                                if (n > kScan && !hit) {   // a long bucket (rare)
                                    for (uint32_t i = st + kScan; i < st + n; i++) {
                                        const uint2 e = tkv[i];
                                        if (e.x == k) { rr = e.y; hit = true; break; }
                                    }
                                }
                                v[j].x = rr;   // the pair (R rid, S rid), kept only if hit
                                hitm |= (hit ? 1u : 0u) << j;
                            }
                            // wait here for the next chunk's loads (issued a probe
                            // ago), not at the next chunk's start, where the wait
                            // would also cover the stores below
                            __builtin_amdgcn_s_waitcnt(0);
                            c = cn;
                            cn = cn < nch ? __builtin_amdgcn_readfirstlane(cg) : nch;
                            wave_emit_round<IPT>(hitm, v, &sh, I.out, erri);
                        }
                        q++;
                    }
                };
                if constexpr (kRange) {
                    if (alg == 2) table_join(std::true_type{});
                    else table_join(std::false_type{});
                } else {
                    table_join(std::false_type{});
                }
                if (to_smj) {   // sh.fill was reset above (hash), or nothing was written (sort-merge)
                    alg = 2;
                    run_sort = true;
                    continue;
                }
            }
            break;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            if (sh.fill) atomicAdd(&out_fill[w.inst], sh.fill);
            if (sh.fill < sh.res) {
                // hole [base + fill, base + res) for the compaction step
                const uint32_t h = atomicAdd(&ctrl[2], 1u);
                if (h < max_holes) holes[h] = make_uint4(w.inst, sh.base + sh.fill, sh.res - sh.fill, 0u);
            }
            // optionally, where the item's pairs went: (instance, partition,
            // first position, count), for results grouped by partition
            if (ext) ext[it] = make_uint4(w.inst, items[it].part, sh.base, sh.fill);   // (read here: nothing held across the item)
            if (!advanced) {   // (the hash path took the next item at its first sub-table)
                sh.item = sh.item2;
                sh.item2 = atomicAdd(&ctrl[1], 1u);
            }
        }
    }
}

// Copy segments of pairs: seg = (src offset, dst offset, length) in pairs.
// ---------------------------------------------------------------------------
// No-partition hash join (packetwise_no_partitioned_hash_join_build.c /
// _probe.c: one table over the whole R, probed by every S tuple).  Here the
// table is one open-addressing array in HBM per instance, 2^log2cap >= 2|R|
// slots of (rid << 32 | key), linear probing from a multiplicative hash;
// key 0 marks an empty slot (key 0 never joins).  Small tables stay in L2
// and the Infinity Cache, which is where this path pays: no partition pass.
// Grid (blocks over tuples, jobs).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t nphj_slot(uint32_t key, uint32_t log2cap)
{
    return (key * 0x9E3779B1u) >> (32 - log2cap);
}

// the job owning block b of a launch: the last job whose first block <= b
__device__ __forceinline__ uint32_t nphj_job_of(const NphjJob *jobs, uint32_t njobs, uint32_t b, bool probe)
{
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t m = (lo + hi) >> 1;
        if ((probe ? jobs[m].first_probe : jobs[m].first_build) <= b) lo = m; else hi = m;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_nphj_clear(unsigned long long *__restrict__ t, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        t[i] = 0ull;
}

__global__ __launch_bounds__(256) void k_nphj_build(const InstDesc *__restrict__ insts, const NphjJob *__restrict__ jobs,
                                                    uint32_t njobs, unsigned long long *__restrict__ tables,
                                                    uint32_t *__restrict__ err)
{
    const NphjJob job = jobs[nphj_job_of(jobs, njobs, blockIdx.x, false)];
    const InstDesc &I = insts[job.inst];
    unsigned long long *tab = tables + job.table;
    const uint32_t mask = (1u << job.log2cap) - 1u;
    const uint64_t b0 = (uint64_t)(blockIdx.x - job.first_build) * kNphjBuildTuples;
    for (uint32_t u = 0; u < kNphjBuildTuples / 256; u++) {
        const uint64_t i = b0 + u * 256 + threadIdx.x;
        if (i >= I.nR) break;
        const uint2 t = gld(I.R, i);
        if (t.x == kEmptyKey) continue;
        const unsigned long long v = ((unsigned long long)t.y << 32) | t.x;
        // at most |R| of the >= 2|R| slots are taken: an empty slot is found
        for (uint32_t h = nphj_slot(t.x, job.log2cap);; h = (h + 1u) & mask) {
            const unsigned long long old = atomicCAS(&tab[h], 0ull, v);
            if (old == 0ull) break;
            if ((uint32_t)old == t.x) {   // a2a_srj_build.c:265-271: duplicate R key
                atomicOr(&err[job.inst], kErrDuplicateKey);
                break;
            }
        }
    }
}

// Each block probes kNphjProbeTuples S tuples and reserves their output
// with one atomic (a block scan of the threads' hit counts gives each thread
// its place): per-wave reservations would put one atomic per 64 tuples on
// the same counter.
__global__ __launch_bounds__(256) void k_nphj_probe(const InstDesc *__restrict__ insts, const NphjJob *__restrict__ jobs,
                                                    uint32_t njobs, const unsigned long long *__restrict__ tables,
                                                    uint32_t *__restrict__ out_lo, uint32_t *__restrict__ out_fill)
{
    __shared__ uint32_t cnt[256], wtmp[256 / kWave + 1], base;
    const NphjJob job = jobs[nphj_job_of(jobs, njobs, blockIdx.x, true)];
    const InstDesc &I = insts[job.inst];
    const unsigned long long *tab = tables + job.table;
    const uint32_t mask = (1u << job.log2cap) - 1u;
    const uint64_t b0 = (uint64_t)(blockIdx.x - job.first_probe) * kNphjProbeTuples;
    uint2 t[kNphjPPT];
#pragma unroll
    for (uint32_t u = 0; u < kNphjPPT; u++) {
        const uint64_t i = b0 + u * 256 + threadIdx.x;
        t[u] = i < I.nS ? gld(I.S, i) : make_uint2(kEmptyKey, 0u);
    }
    uint32_t hits = 0, n = 0;
#pragma unroll
    for (uint32_t u = 0; u < kNphjPPT; u++) {
        if (t[u].x == kEmptyKey) continue;
        for (uint32_t h = nphj_slot(t[u].x, job.log2cap);; h = (h + 1u) & mask) {
            const unsigned long long v = tab[h];
            if (v == 0ull) break;
            if ((uint32_t)v == t[u].x) {
                hits |= 1u << u;
                t[u].x = (uint32_t)(v >> 32);   // the R rid replaces the key
                n++;
                break;
            }
        }
    }
    cnt[threadIdx.x] = n;
    const uint32_t total = block_exclusive_scan(cnt, 256, wtmp);   // cnt[t] = thread t's first place
    if (total == 0) return;   // the same in every thread
    if (threadIdx.x == 0) base = atomicAdd(&out_lo[job.inst], total);
    __syncthreads();
    uint32_t o = base + cnt[threadIdx.x];
#pragma unroll
    for (uint32_t u = 0; u < kNphjPPT; u++)
        if (hits & (1u << u)) gst(I.out, o++, t[u]);   // (R rid, S rid)
    if (threadIdx.x == 0) atomicAdd(&out_fill[job.inst], total);
}

__global__ void k_gather_segments(const uint2 *__restrict__ src, uint2 *__restrict__ dst, const uint4 *__restrict__ segs,
                                  uint32_t nsegs)
{
    for (uint32_t s = blockIdx.x; s < nsegs; s += gridDim.x) {
        const uint4 g = segs[s];
        for (uint32_t i = threadIdx.x; i < g.z; i += blockDim.x) dst[g.y + i] = src[g.x + i];
    }
}

// ---------------------------------------------------------------------------
// Order-independent checksum of a pair array (sum and xor of a 64-bit mix of
// r_rid << 32 | s_rid): lets a caller compare two result multisets without
// copying or sorting them.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_pairs_checksum(const uint2 *__restrict__ p, uint64_t n,
                                                        unsigned long long *__restrict__ acc)
{
    uint64_t s = 0, x = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint2 v = p[i];
        const uint64_t h = mix64(((uint64_t)v.x << 32) | v.y);
        s += h;
        x ^= h;
    }
    for (uint32_t off = kWave / 2; off > 0; off >>= 1) {
        s += __shfl_xor(s, off, kWave);
        x ^= __shfl_xor(x, off, kWave);
    }
    if (lane_id() == 0) {
        atomicAdd(&acc[0], (unsigned long long)s);
        atomicXor(&acc[1], (unsigned long long)x);
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
// Allow `bytes` of dynamic LDS for kernel fn on the current device.  The
// attribute is set once per (device, kernel) and raised only when a launch
// asks for more: it is a runtime call of several microseconds, which every
// launch would otherwise pay while the GPU waits.
static hipError_t set_lds(const void *fn, size_t bytes)
{
    if (bytes <= 65536) return hipSuccess;
    static std::mutex mu;
    static std::vector<std::pair<std::pair<int, const void *>, size_t>> done;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(mu);
    for (auto &d : done)
        if (d.first.first == dev && d.first.second == fn) {
            if (d.second >= bytes) return hipSuccess;
            e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
            if (e == hipSuccess) d.second = bytes;
            return e;
        }
    e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess) done.push_back({{dev, fn}, bytes});
    return e;
}

size_t lds_hist(uint32_t max_bits)
{
    return ((size_t(1) << max_bits) + kHeavySlots) * 4;
}

size_t lds_scatter(uint32_t max_bits, int ipt)
{
    const size_t F = size_t(1) << max_bits, R = (size_t)kScatterNT * (size_t)ipt;
    return (size_t)kHeavyWords * 4 + R * 8 + 2 * (F + 2) * 4 +
           (kScatterNT / kWave + 1) * 4 + 16;
}

size_t lds_join()
{
    // the hash table and its bucket ends, then the sort-merge list of long buckets
    size_t hash = kBigListOff + (size_t)kBigList * 2;
    size_t smj = ((size_t)kSortPad + kSortTail) * 8;
    size_t nl = 2 * kNLChunk * 4;
    size_t m = hash > smj ? hash : smj;
    return m > nl ? m : nl;
}

uint32_t join_table_cap() { return kTableCap; }

hipError_t launch_hist(const InstDesc *insts, const TileDesc *tiles, uint32_t ntiles, uint32_t *counts,
                       uint32_t max_bits, hipStream_t st)
{
    if (!ntiles) return hipSuccess;
    size_t l = lds_hist(max_bits);
    hipError_t e = set_lds((const void *)k_hist, l);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_hist, dim3(ntiles), dim3(1024), l, st, insts, tiles, ntiles, counts);
    return hipGetLastError();
}

hipError_t launch_scan(const ScanSeg *segs, uint32_t nsegs, uint32_t *counts, uint32_t *pstart, uint32_t max_bits,
                       hipStream_t st)
{
    if (!nsegs) return hipSuccess;
    size_t l = (size_t(1) << max_bits) * 4;
    hipError_t e = set_lds((const void *)k_scan, l);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan, dim3(nsegs), dim3(1024), l, st, segs, counts, pstart);
    return hipGetLastError();
}

hipError_t launch_scatter(const InstDesc *insts, const TileDesc *tiles, uint32_t ntiles, uint32_t *counts,
                          uint32_t *out_hi, uint32_t *err, uint32_t *tq, uint32_t max_bits, int ipt, uint32_t grid,
                          hipStream_t st)
{
    if (!ntiles) return hipSuccess;
    grid = (grid + 7u) & ~7u;   // whole XCD groups
    size_t l = lds_scatter(max_bits, ipt);
    hipError_t e;
    if (ipt == 12) {
        e = set_lds((const void *)k_scatter<12>, l);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_scatter<12>, dim3(grid), dim3(kScatterNT), l, st, insts, tiles, ntiles, max_bits, counts,
                           out_hi, err, tq);
    } else if (ipt == 10) {
        e = set_lds((const void *)k_scatter<10>, l);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_scatter<10>, dim3(grid), dim3(kScatterNT), l, st, insts, tiles, ntiles, max_bits, counts,
                           out_hi, err, tq);
    } else if (ipt == 8) {
        e = set_lds((const void *)k_scatter<8>, l);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_scatter<8>, dim3(grid), dim3(kScatterNT), l, st, insts, tiles, ntiles, max_bits, counts,
                           out_hi, err, tq);
    } else {
        e = set_lds((const void *)k_scatter<4>, l);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_scatter<4>, dim3(grid), dim3(kScatterNT), l, st, insts, tiles, ntiles, max_bits, counts,
                           out_hi, err, tq);
    }
    return hipGetLastError();
}

uint32_t heavy_sample_blocks(uint32_t nS, uint32_t max_blocks, uint32_t *stride)
{
    const uint32_t total = (nS + kSampleBlock - 1) / kSampleBlock;
    // sample at most 1/16 of the blocks (and at most max_blocks of them)
    const uint32_t want = min(max_blocks, (total + 15) / 16);
    uint32_t st = want ? total / want : 1;
    if (st < 1) st = 1;
    if (stride) *stride = st;
    return total ? (total + st - 1) / st : 0;
}

hipError_t launch_heavy_detect(const InstDesc *insts, const HeavyJob *jobs, uint32_t njobs, const uint32_t blocks[2],
                               uint32_t *cand_base, uint32_t *nheavy, uint32_t grid, hipStream_t st)
{
    if (!njobs || !blocks[0]) return hipSuccess;
    size_t l = (2 * kSampleSlots + 2 * kAccSlots) * 4;
    hipError_t e = set_lds((const void *)k_heavy_sample, l);
    if (e != hipSuccess) return e;
    const uint32_t g0 = min(blocks[0], grid), g1 = min(blocks[1], grid);
    hipLaunchKernelGGL(k_heavy_sample, dim3(g0), dim3(1024), l, st, insts, jobs, njobs, cand_base, 0u, blocks[0]);
    if (blocks[1])
        hipLaunchKernelGGL(k_heavy_sample, dim3(g1), dim3(1024), l, st, insts, jobs, njobs, cand_base, 1u, blocks[1]);
    hipLaunchKernelGGL(k_heavy_select, dim3(njobs), dim3(1024), 0, st, insts, jobs, cand_base, nheavy);
    return hipGetLastError();
}

uint32_t cand_words() { return 2 * kCandSlots + 4; }

hipError_t launch_prep(const PrepArgs &a, hipStream_t st)
{
    const uint64_t flat = a.nctrl + a.nheavy + a.ncand + a.nzero + a.ntiles + a.nnphj;
    uint64_t bx = (flat + 255) / 256;
    if (bx < a.ncells) bx = a.ncells;
    if (bx > 65536) bx = 65536;   // (large no-partition tables: 16M threads, as their own fill had)
    if (!bx) return hipSuccess;
    hipLaunchKernelGGL(k_prep, dim3((uint32_t)bx, a.ncells ? 2u : 1u), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_make_items(const InstDesc *insts, uint32_t ninst, const uint32_t *pstart, const CellGroup *cells,
                             const uint32_t *counts, uint32_t smax, WorkItem *items, uint32_t *nitems,
                             uint32_t *spart, uint32_t *rpart, hipStream_t st)
{
    if (!ninst) return hipSuccess;
    hipLaunchKernelGGL(k_make_items, dim3(ninst, (1u << kMaxBits) / kItemsNT), dim3(kItemsNT), 0, st, insts, pstart, cells,
                       counts, smax, items, nitems, spart, rpart);
    return hipGetLastError();
}

hipError_t launch_join(const InstDesc *insts, const WorkItem *items, uint32_t *ctrl,
                       uint32_t *out_lo, uint32_t *out_fill, uint32_t *err, uint4 *holes, uint32_t max_holes,
                       uint4 *ext, uint32_t grid, bool range, hipStream_t st)
{
    size_t l = lds_join();
    const void *k = range ? (const void *)k_join<true> : (const void *)k_join<false>;
    hipError_t e = set_lds(k, l);
    if (e != hipSuccess) return e;
    if (range)
        hipLaunchKernelGGL(k_join<true>, dim3(grid), dim3(kJoinNT), l, st, insts, items, ctrl, out_lo, out_fill, err,
                           holes, max_holes, ext);
    else
        hipLaunchKernelGGL(k_join<false>, dim3(grid), dim3(kJoinNT), l, st, insts, items, ctrl, out_lo, out_fill, err,
                           holes, max_holes, ext);
    return hipGetLastError();
}

hipError_t launch_nphj(const InstDesc *insts, const NphjJob *jobs, uint32_t njobs, uint64_t *tables,
                       uint64_t table_words, uint32_t build_blocks, uint32_t probe_blocks, uint32_t *out_lo,
                       uint32_t *out_fill, uint32_t *err, bool clear, hipStream_t st)
{
    if (!njobs) return hipSuccess;
    // a fill kernel rather than a memset node, so that the sequence stays
    // capturable into a graph (the batch path clears the tables in k_prep)
    auto *t = reinterpret_cast<unsigned long long *>(tables);
    const uint64_t fb = std::min<uint64_t>((table_words + 255) / 256, 65536);
    if (clear) hipLaunchKernelGGL(k_nphj_clear, dim3((uint32_t)fb), dim3(256), 0, st, t, table_words);
    if (build_blocks)
        hipLaunchKernelGGL(k_nphj_build, dim3(build_blocks), dim3(256), 0, st, insts, jobs, njobs, t, err);
    if (probe_blocks)
        hipLaunchKernelGGL(k_nphj_probe, dim3(probe_blocks), dim3(256), 0, st, insts, jobs, njobs,
                           (const unsigned long long *)t, out_lo, out_fill);
    return hipGetLastError();
}

hipError_t launch_pairs_checksum(const uint2 *p, uint64_t n, unsigned long long *acc, hipStream_t st)
{
    hipError_t e = hipMemsetAsync(acc, 0, 16, st);
    if (e != hipSuccess || !n) return e;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_pairs_checksum, dim3((uint32_t)blocks), dim3(256), 0, st, p, n, acc);
    return hipGetLastError();
}

hipError_t launch_gather_segments(const uint2 *src, uint2 *dst, const uint4 *segs, uint32_t nsegs, hipStream_t st)
{
    if (!nsegs) return hipSuccess;
    uint32_t grid = nsegs < 4096 ? nsegs : 4096;
    hipLaunchKernelGGL(k_gather_segments, dim3(grid), dim3(256), 0, st, src, dst, segs, nsegs);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The no-allocation join: the no-partition hash join of one instance with
// its descriptors, control words and table in a static device array, for
// callers that cannot allocate.  Layout of `scratch` (8-byte words):
//   [0, kNaHead)       InstDesc, NphjJob, then the control words out_lo,
//                      out_fill and err (3 x uint32)
//   [kNaHead, ...)     the open-addressing table, 2^log2cap slots
// Host code reaches the array only through the address hipGetSymbolAddress
// returns on each device (scratch_on_device); its name in host code is not
// a device address.
// ---------------------------------------------------------------------------
constexpr uint64_t kNaHead = 64;   // 512 bytes of descriptors and control words
constexpr uint32_t kNaMaxLog2Cap = 21;   // >= 2 |R| slots for |R| <= kNoAllocMaxR
static_assert((uint64_t(1) << kNaMaxLog2Cap) >= 2 * kNoAllocMaxR, "the table holds 2 |R| slots");
static_assert(sizeof(InstDesc) + sizeof(NphjJob) + 3 * 4 <= kNaHead * 8, "descriptors fit the head");
constexpr uint64_t kNaWords = kNaHead + (uint64_t(1) << kNaMaxLog2Cap);
static __device__ __align__(256) unsigned long long scratch[kNaWords];

// Write the descriptors (kernel arguments) into the head and zero the
// control words: no host-to-device copy is needed.
__global__ void k_noalloc_setup(InstDesc d, NphjJob j, unsigned long long *head)
{
    if (threadIdx.x != 0) return;
    *reinterpret_cast<InstDesc *>(head) = d;
    *reinterpret_cast<NphjJob *>(reinterpret_cast<char *>(head) + sizeof(InstDesc)) = j;
    uint32_t *ctrl = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(head) + sizeof(InstDesc) + sizeof(NphjJob));
    ctrl[0] = ctrl[1] = ctrl[2] = 0u;
}

// The scratch array's device address on the current device, looked up once
// per device (each device holds its own copy of the array).
static std::mutex g_na_mu;
static void *g_na_ptr[kNoAllocMaxDevices] = {};

hipError_t scratch_on_device(void **ptr)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= (int)kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_na_ptr[dev] = p;
    }
    *ptr = g_na_ptr[dev];
    return hipSuccess;
}

hipError_t run_nphj_noalloc(const uint2 *R, uint32_t nR, const uint2 *S, uint32_t nS, uint2 *out, uint32_t ctrl_out[3],
                            hipStream_t st)
{
    if (nR > kNoAllocMaxR) return hipErrorInvalidValue;
    void *base = nullptr;
    hipError_t e = scratch_on_device(&base);
    if (e != hipSuccess) return e;
    auto *head = reinterpret_cast<unsigned long long *>(base);
    InstDesc d{};
    d.R = R;
    d.S = S;
    d.out = out;
    d.nR = nR;
    d.nS = nS;
    d.alg = 1;
    NphjJob j{};
    j.inst = 0;
    j.log2cap = 10;
    while ((1ull << j.log2cap) < 2ull * nR) j.log2cap++;
    j.table = 0;
    j.first_build = 0;
    j.first_probe = 0;
    const InstDesc *d_inst = reinterpret_cast<const InstDesc *>(head);
    const NphjJob *d_job = reinterpret_cast<const NphjJob *>(reinterpret_cast<char *>(head) + sizeof(InstDesc));
    uint32_t *ctrl = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(head) + sizeof(InstDesc) + sizeof(NphjJob));
    hipLaunchKernelGGL(k_noalloc_setup, dim3(1), dim3(64), 0, st, d, j, head);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint32_t bb = (nR + kNphjBuildTuples - 1) / kNphjBuildTuples;
    const uint32_t pb = (uint32_t)(((uint64_t)nS + kNphjProbeTuples - 1) / kNphjProbeTuples);
    e = launch_nphj(d_inst, d_job, 1, reinterpret_cast<uint64_t *>(head + kNaHead), uint64_t(1) << j.log2cap, bb, pb,
                    ctrl, ctrl + 1, ctrl + 2, true, st);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(ctrl_out, ctrl, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(st);
}

}  // namespace spj
