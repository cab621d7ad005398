// spj_capi.cpp -- extern "C" surface of libspidjoin.so (include/spidjoin.h).
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "spidjoin.h"
#include "spj_capi_internal.hpp"
#include "spj_stage.hpp"
#include "spj_engine.hpp"
#include "spj_kernels.hpp"
#include "spj_tuner.hpp"

namespace spj {

static thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    if (e == hipErrorOutOfMemory) return fail(SPJ_ENOMEM, "%s: %s", what, hipGetErrorString(e));
    return fail(SPJ_EHIP, "%s: %s", what, hipGetErrorString(e));
}

int err_from_flags(uint32_t f, size_t inst)
{
    if (f & kErrDuplicateKey) return fail(SPJ_EDUPKEY, "instance %zu: duplicate key on the R (build) side", inst);
    if (f & kErrOutOverflow) return fail(SPJ_EOVERFLOW, "instance %zu: output capacity exceeded", inst);
    return SPJ_OK;
}

Resolved resolve(const spj_config_t *cfg, uint64_t nR, uint64_t nS, double skew_hint)
{
    Resolved r = tuned_params((uint32_t)nR, (uint32_t)nS, skew_hint);
    if (!cfg) return r;
    if (cfg->algorithm >= 0) r.alg = (uint32_t)cfg->algorithm;
    if (cfg->radix_bits >= 0) r.bits = (uint32_t)cfg->radix_bits;
    if (cfg->table_log2 >= 0) r.table_log2 = (uint32_t)cfg->table_log2;
    if (cfg->heavy >= 0) r.heavy = (uint32_t)cfg->heavy;
    if (cfg->heavy_min_count >= 0) r.heavy_min = (uint32_t)cfg->heavy_min_count;
    if (cfg->scatter_ipt >= 0) r.ipt = (uint32_t)cfg->scatter_ipt;
    if (cfg->tile_tuples > 0) r.tile = (uint32_t)cfg->tile_tuples;
    if (cfg->items_s_max > 0) r.smax = (uint32_t)cfg->items_s_max;
    if (cfg->part_layout >= 0) r.cells = (uint32_t)cfg->part_layout;
    if (cfg->cell_capacity_pct >= 0) r.cell_pct = (uint32_t)cfg->cell_capacity_pct;
    if (cfg->radix_passes >= 0) r.passes = (uint32_t)cfg->radix_passes;
    return r;
}

// AUTO's choice of the no-partition join, measured on MI355X (device
// relations, Zipf 0.5, ms per call, partitioned / no-partition):
//   |R|     S:R = 4          S:R = 64         S:R = 256
//   2^18    0.301 / 0.090    0.460 / 0.238    0.804 / 0.783
//   2^19    0.260 / 0.130    0.577 / 0.515    1.470 / 1.724
//   2^20    0.301 / 0.200    0.927 / 1.152    2.662 / 4.163
//   2^21    0.369 / 0.310    1.554 / 2.507
//   2^22    0.365 / 0.645    2.695 / 5.810
// Without a partition pass the join wins while its table (16 B per R
// tuple) stays cached and the partition pipeline's fixed cost dominates;
// each S tuple's random table read costs more the larger R is.  A single
// call takes it for |R| <= 2^21 and |R| |S| <= 2^44.  A batch is judged by
// its totals, with |R| <= 2^20: the tables of a large batch do not stay
// cached (1024 mixed instances: 6.7 ms partitioned, 16.0 ms without).
constexpr uint64_t kAutoNphjMaxR = uint64_t(1) << 21;
constexpr uint64_t kAutoNphjMaxBatchR = uint64_t(1) << 20;
constexpr uint32_t kAutoNphjMaxLog2RS = 44;
constexpr uint64_t kNphjMaxR = uint64_t(1) << 30;   // its table's slot index is 32-bit

static bool auto_sizes(uint64_t r, uint64_t s, uint64_t max_r)
{
    if (r > max_r) return false;
    // |R| |S| <= 2^44 without overflow: |S| <= 2^44 / |R|
    return r == 0 || s <= (uint64_t(1) << kAutoNphjMaxLog2RS) / r;
}

static bool auto_no_partition(const spj_config_t *cfg, uint64_t r, uint64_t s, uint64_t max_r)
{
    if (!auto_sizes(r, s, max_r)) return false;
    if (!cfg) return true;
    return cfg->algorithm < 0 && cfg->radix_bits < 0 && cfg->table_log2 < 0 && cfg->heavy < 0 &&
           cfg->heavy_min_count < 0 && cfg->scatter_ipt < 0 && cfg->tile_tuples <= 0 && cfg->items_s_max <= 0 &&
           cfg->part_layout < 0 && cfg->cell_capacity_pct < 0 && cfg->radix_passes < 0;
}

// AUTO's no-partition choice for one instance: the rule above, unless the
// tuner holds an entry for the instance's bucket, which then decides (its
// entry may itself be the no-partition join).
bool auto_no_partition_single(const spj_config_t *cfg, uint64_t nR, uint64_t nS)
{
    return auto_no_partition(cfg, nR, nS, kAutoNphjMaxR) && !tuner_has_entry((uint32_t)nR, (uint32_t)nS, -1.0);
}

int check_resolved(const Resolved &r, uint64_t nR)
{
    if (r.alg == SPJ_ALG_NPHJ && nR > kNphjMaxR)
        return fail(SPJ_EINVAL, "the no-partition join takes at most 2^30 R tuples (its table has 2^k >= 2|R| "
                                "slots), got %llu",
                    (unsigned long long)nR);
    if (r.bits > kMaxBits) return fail(SPJ_EINVAL, "radix_bits %u > %u", r.bits, kMaxBits);
    if (r.table_log2 < 6 || r.table_log2 > 13) return fail(SPJ_EINVAL, "table_log2 %u outside 6..13", r.table_log2);
    if (r.alg > 3) return fail(SPJ_EINVAL, "algorithm %u", r.alg);
    if (!(r.ipt == 4 || r.ipt == 8 || r.ipt == 10 || r.ipt == 12))
        return fail(SPJ_EINVAL, "scatter_ipt %u (4, 8, 10 or 12)", r.ipt);
    if (r.tile < 512) return fail(SPJ_EINVAL, "tile_tuples %u < 512", r.tile);
    if (r.smax < 512) return fail(SPJ_EINVAL, "items_s_max %u < 512", r.smax);
    if (r.cells > 2) return fail(SPJ_EINVAL, "part_layout %u (0 counted, 1 cells)", r.cells);
    if (r.cell_pct > 200) return fail(SPJ_EINVAL, "cell_capacity_pct %u > 200", r.cell_pct);
    if (r.passes < 1 || r.passes > 2) return fail(SPJ_EINVAL, "radix_passes %u (1 or 2)", r.passes);
    if (r.passes == 2 && (r.alg == SPJ_ALG_NPHJ || r.bits == 0))
        return fail(SPJ_EINVAL, "two radix passes need a partitioned join and radix_bits >= 1");
    return SPJ_OK;
}

static std::mutex g_ctx_mu;
static Context *g_ctx = nullptr;

Context &default_context()
{
    if (!g_ctx) g_ctx = new Context();
    return *g_ctx;
}

std::mutex &default_context_mutex() { return g_ctx_mu; }

void fill_times(const PhaseTimes &t, spj_phase_times_t *o)
{
    if (!o) return;
    o->heavy_ms = t.heavy;
    o->hist_ms = t.hist;
    o->scan_ms = t.scan;
    o->scatter_ms = t.scatter;
    o->items_ms = t.items;
    o->join_ms = t.join;
    o->total_ms = t.total;
    o->heavy_keys = t.heavy_keys;
    o->r_partitioned = t.r_parted;
    o->s_partitioned = t.s_parted;
    o->s_cells = t.s_cells;
    o->r_cells = t.r_cells;
    o->passes = t.passes;
}

int validate_sizes(uint64_t nR, uint64_t nS)
{
    if (nR >= 0xFFFFFFFFull || nS >= 0xFFFFFFFFull)
        return fail(SPJ_EINVAL, "relation sizes must be below 2^32 - 1 tuples (got %llu, %llu)",
                    (unsigned long long)nR, (unsigned long long)nS);
    return SPJ_OK;
}

}  // namespace spj

using namespace spj;

extern "C" {

void spj_config_default(spj_config_t *cfg)
{
    if (!cfg) return;
    cfg->algorithm = SPJ_ALG_AUTO;
    cfg->radix_bits = -1;
    cfg->table_log2 = -1;
    cfg->heavy = -1;
    cfg->heavy_min_count = -1;
    cfg->scatter_ipt = -1;
    cfg->tile_tuples = -1;
    cfg->items_s_max = -1;
    cfg->part_layout = -1;
    cfg->cell_capacity_pct = -1;
    cfg->radix_passes = -1;
}

const char *spj_last_error(void) { return g_err.c_str(); }

int spj_init(int device)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    if (device < 0 || device >= n) return fail(SPJ_EINVAL, "device %d of %d", device, n);
    e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    (void)device_info();
    void *scratch = nullptr;   // the no-allocation join's array on this device
    e = scratch_on_device(&scratch);
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    return SPJ_OK;
}

int spj_device_info(int *cus, uint64_t *lds, char *arch, int arch_len)
{
    const DeviceInfo &d = device_info();
    if (d.cus == 0) return fail(SPJ_EHIP, "no HIP device");
    if (cus) *cus = d.cus;
    if (lds) *lds = d.lds_per_block;
    if (arch && arch_len > 0) {
        strncpy(arch, d.arch.c_str(), (size_t)arch_len - 1);
        arch[arch_len - 1] = 0;
    }
    return SPJ_OK;
}

int spj_dev_alloc(void **p, uint64_t bytes)
{
    if (!p) return fail(SPJ_EINVAL, "null pointer");
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    return e == hipSuccess ? SPJ_OK : hip_fail(e, "hipMalloc");
}

int spj_dev_free(void *p)
{
    hipError_t e = hipFree(p);
    return e == hipSuccess ? SPJ_OK : hip_fail(e, "hipFree");
}

int spj_h2d(void *dst, const void *src, uint64_t bytes)
{
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    return e == hipSuccess ? SPJ_OK : hip_fail(e, "hipMemcpy H2D");
}

int spj_d2h(void *dst, const void *src, uint64_t bytes)
{
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
    return e == hipSuccess ? SPJ_OK : hip_fail(e, "hipMemcpy D2H");
}

int spj_dev_memset(void *p, int v, uint64_t bytes)
{
    hipError_t e = hipMemset(p, v, bytes);
    return e == hipSuccess ? SPJ_OK : hip_fail(e, "hipMemset");
}

int spj_d2d(void *dst, const void *src, uint64_t bytes, float *ms)
{
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipEventRecord(a, 0);
    if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, 0);
    if (e == hipSuccess) e = hipEventRecord(b, 0);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess && ms) e = hipEventElapsedTime(ms, a, b);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    return e == hipSuccess ? SPJ_OK : hip_fail(e, "hipMemcpy D2D");
}

int spj_pairs_checksum_device(const spj_pair_t *pairs, uint64_t n, uint64_t *sum, uint64_t *xr)
{
    if ((!pairs && n) || !sum || !xr) return fail(SPJ_EINVAL, "null argument");
    unsigned long long *acc = nullptr, h[2] = {0, 0};
    hipError_t e = hipMalloc(&acc, 16);
    if (e == hipSuccess) e = launch_pairs_checksum(reinterpret_cast<const uint2 *>(pairs), n, acc, 0);
    if (e == hipSuccess) e = hipMemcpy(h, acc, 16, hipMemcpyDeviceToHost);
    if (acc) (void)hipFree(acc);
    if (e != hipSuccess) return hip_fail(e, "checksum");
    *sum = h[0];
    *xr = h[1];
    return SPJ_OK;
}

int spj_graph_stats(uint64_t *captures, uint64_t *replays)
{
    if (!captures || !replays) return fail(SPJ_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(default_context_mutex());
    Context &c = default_context();
    *captures = c.graph_captures();
    *replays = c.graph_replays();
    return SPJ_OK;
}

int spj_dev_sync(void)
{
    hipError_t e = hipDeviceSynchronize();
    return e == hipSuccess ? SPJ_OK : hip_fail(e, "hipDeviceSynchronize");
}

}  // extern "C"

namespace spj {

// The second pass's context: its buffers stay apart from the first pass's,
// whose partitions it reads.  Used under the default context's lock.
static Context &second_pass_context()
{
    static Context *c = nullptr;
    if (!c) c = new Context();
    return *c;
}

// A join in two radix passes.  The first partitions R and S by s0.cfg.bits
// of SuperFastHash into contiguous partitions; heavy keys (if on) are
// joined in its S pass, their pairs at the back of out.  The second is a
// batch with one instance per first-pass partition: partition p of R and S,
// split by the next bits (as many as make its R partitions fit one LDS
// table) and joined, its pairs written from out + (start of S partition p).
// Partitions are then packed: every instance's pairs, in partition order,
// then the heavy pairs.  ref: the reference partitions twice too, globally
// across DPUs (glb_partition_count.c) and locally inside each
// (a2a_packetwise_srj_local_partition.c:323-350).
static hipError_t run_two_pass(const InstSpec &s0, uint32_t &count, uint32_t &err, PhaseTimes *times)
{
    Context &c1 = default_context(), &c2 = second_pass_context();
    // The first pass leaves a relation's partitions in cells (no histogram
    // pass; the second pass's instances read a partition as its groups'
    // segments) unless the caller fixed the layout or the segments would be
    // short: a segment starts a scatter tile of its own, and segments of a
    // few thousand tuples leave the second pass's rounds mostly empty (2^26
    // x 2^26 in two passes, 2K-tuple segments: 2.65 ms against 2.23 counted).
    constexpr uint64_t kMinSegment = 4096;
    InstSpec p1 = s0;
    p1.partition_only = true;
    p1.r_contiguous = ((uint64_t)s0.nR >> s0.cfg.bits) / kScatterGroups < kMinSegment;
    p1.s_contiguous = ((uint64_t)s0.nS >> s0.cfg.bits) / kScatterGroups < kMinSegment;
    std::vector<uint32_t> cnt, errs;
    PhaseTimes t1, t2;
    hipError_t e = c1.run({p1}, cnt, errs, times ? &t1 : nullptr);
    if (e != hipSuccess) return e;
    const uint32_t heavy_pairs = cnt[0];
    err = errs[0];
    const uint2 *rp = nullptr, *sp = nullptr;
    uint32_t ngr = 0, ngs = 0;
    std::vector<uint2> sr, ss;
    e = c1.partition_segments(0, &rp, &sp, ngr, sr, ngs, ss);
    if (e != hipSuccess) return e;
    const uint32_t F1 = 1u << s0.cfg.bits;
    const uint64_t seg_s = ((uint64_t)s0.nS >> s0.cfg.bits) / ngs;   // mean S segment
    std::vector<InstSpec> sub;
    std::vector<uint32_t> where;   // each instance's output offset: the S tuples of the partitions before it
    uint64_t r_parted = 0, s_parted = 0;
    for (uint32_t p = 0; p < F1; p++) {
        InstSpec q;
        uint64_t nr = 0, ns = 0;
        for (uint32_t k = 0; k < ngr; k++)
            if (sr[(size_t)p * ngr + k].y) {
                q.rseg.push_back(sr[(size_t)p * ngr + k]);
                nr += sr[(size_t)p * ngr + k].y;
            }
        for (uint32_t k = 0; k < ngs; k++)
            if (ss[(size_t)p * ngs + k].y) {
                q.sseg.push_back(ss[(size_t)p * ngs + k]);
                ns += ss[(size_t)p * ngs + k].y;
            }
        const uint64_t at = s_parted;
        r_parted += nr;
        s_parted += ns;
        if (!nr || !ns) continue;   // no pairs: an S partition without R tuples matches nothing
        q.R = rp;
        q.nR = (uint32_t)nr;
        q.S = sp;
        q.nS = (uint32_t)ns;
        q.out = s0.out + at;
        q.shift = s0.cfg.bits;
        Resolved c = default_params(nr, ns);
        c.bits = std::min(c.bits, kMaxDigitBits - s0.cfg.bits);
        c.alg = s0.cfg.alg;
        c.table_log2 = s0.cfg.table_log2;
        // rounds of its own size: a segment (a group's share of a
        // first-pass partition) ends in a partly filled round, and the
        // second pass's few digits make small rounds cheap.  Measured at
        // 2^28 x 2^29 (16K-tuple S segments; 12 / 10 / 8 tuples per thread):
        // 9.59 / 9.40 / 9.76 ms; at 2^28 x 2^30 (32K): 17.03 / 15.92 / 15.36.
        c.ipt = seg_s >= 24576 ? 8u : 10u;
        c.smax = s0.cfg.smax;   // one value per batch
        // cells (no histogram pass) unless the caller fixed the layout: the
        // first pass took the keys hot enough to overflow one.  Cells of
        // twice the mean fill (plus the usual 8 sigma + 64) unless the caller
        // sized them: a partition here is small (about 2K tuples a cell), so
        // keys that the heavy detection leaves alone (Zipf 0.5 at 2^28: 8K
        // occurrences of the hottest) still overflow cells of 1.5 mean +
        // 8 sigma, and a redo of the whole batch counted costs more than the
        // space.  (Twice the mean alone is too tight for a group that gets
        // only a short segment's tail.)
        c.cells = s0.cfg.cells == 2 ? 1u : s0.cfg.cells;
        c.cell_pct = s0.cfg.cell_pct;
        c.cell_mean_x2 = 1;
        c.heavy = 0;   // taken by the first pass
        c.passes = 1;
        q.cfg = c;
        sub.push_back(std::move(q));
        where.push_back((uint32_t)at);
    }
    uint64_t total = heavy_pairs;
    if (!sub.empty()) {
        e = c2.run(sub, cnt, errs, times ? &t2 : nullptr);
        if (e != hipSuccess) return e;
        for (size_t k = 0; k < sub.size(); k++) {
            err |= errs[k];
            total += cnt[k];
        }
    }
    // pack: instance k's pairs at where[k], then the heavy pairs from the back
    std::vector<uint4> segs;
    uint64_t dst = 0;
    bool dense = true;
    for (size_t k = 0; k < sub.size(); k++) {
        if (!cnt[k]) continue;
        dense = dense && where[k] == dst;
        segs.push_back(make_uint4(where[k], (uint32_t)dst, cnt[k], 0));
        dst += cnt[k];
    }
    if (heavy_pairs) {
        const uint32_t hsrc = s0.nS - heavy_pairs;
        dense = dense && hsrc == dst;
        segs.push_back(make_uint4(hsrc, (uint32_t)dst, heavy_pairs, 0));
    }
    if (!dense && !err) {
        e = c2.gather(s0.out, segs, total);
        if (e != hipSuccess) return e;
    }
    count = (uint32_t)total;
    if (times) {
        *times = t1;
        times->hist += t2.hist;
        times->scan += t2.scan;
        times->scatter += t2.scatter;
        times->items = t2.items;
        times->join = t2.join;
        times->total = t1.total + t2.total;
        times->r_parted = r_parted;
        times->s_parted = s_parted;
        times->passes = 2;
    }
    return hipSuccess;
}

hipError_t run_one(const InstSpec &s, uint32_t &count, uint32_t &err, PhaseTimes *times)
{
    if (s.cfg.passes == 2) return run_two_pass(s, count, err, times);
    std::vector<uint32_t> counts, errs;
    const hipError_t e = default_context().run({s}, counts, errs, times);
    if (e != hipSuccess) return e;
    count = counts[0];
    err = errs[0];
    return hipSuccess;
}

}  // namespace spj

extern "C" {

int spj_join_device(const spj_tuple_t *R, uint64_t nR, const spj_tuple_t *S, uint64_t nS, spj_pair_t *out,
                    uint64_t out_cap, uint64_t *out_count, const spj_config_t *cfg, spj_phase_times_t *times)
{
    if ((!R && nR) || (!S && nS) || (!out && nS) || !out_count) return fail(SPJ_EINVAL, "null argument");
    if (out_cap < nS) return fail(SPJ_EINVAL, "out_cap %llu < |S| %llu", (unsigned long long)out_cap, (unsigned long long)nS);
    int rc = validate_sizes(nR, nS);
    if (rc) return rc;
    InstSpec s;
    s.R = reinterpret_cast<const uint2 *>(R);
    s.S = reinterpret_cast<const uint2 *>(S);
    s.out = reinterpret_cast<uint2 *>(out);
    s.nR = (uint32_t)nR;
    s.nS = (uint32_t)nS;
    s.cfg = resolve(cfg, nR, nS, -1.0);
    if (auto_no_partition_single(cfg, nR, nS)) s.cfg.alg = SPJ_ALG_NPHJ;
    rc = check_resolved(s.cfg, nR);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(default_context_mutex());
    PhaseTimes t;
    uint32_t count = 0, err = 0;
    const hipError_t e = run_one(s, count, err, times ? &t : nullptr);
    if (e != hipSuccess) return hip_fail(e, s.cfg.passes == 2 ? "join (two passes)" : "join");
    fill_times(t, times);
    *out_count = count;
    return err_from_flags(err, 0);
}

static_assert(SPJ_NOALLOC_MAX_R == kNoAllocMaxR, "the header states the kernels' limit");

int spj_join_device_noalloc(const spj_tuple_t *R, uint64_t nR, const spj_tuple_t *S, uint64_t nS, spj_pair_t *out,
                            uint64_t out_cap, uint64_t *out_count)
{
    if ((!R && nR) || (!S && nS) || (!out && nS) || !out_count) return fail(SPJ_EINVAL, "null argument");
    if (out_cap < nS) return fail(SPJ_EINVAL, "out_cap %llu < |S| %llu", (unsigned long long)out_cap, (unsigned long long)nS);
    if (nR > kNoAllocMaxR)
        return fail(SPJ_EINVAL, "spj_join_device_noalloc: |R| = %llu exceeds its table's limit of %u tuples",
                    (unsigned long long)nR, kNoAllocMaxR);
    int rc = validate_sizes(nR, nS);
    if (rc) return rc;
    // calls in this process take turns on the device's array
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    uint32_t ctrl[3] = {0, 0, 0};
    hipError_t e = run_nphj_noalloc(reinterpret_cast<const uint2 *>(R), (uint32_t)nR, reinterpret_cast<const uint2 *>(S),
                                    (uint32_t)nS, reinterpret_cast<uint2 *>(out), ctrl, nullptr);
    if (e != hipSuccess) return hip_fail(e, "no-allocation join");
    *out_count = ctrl[1];
    return err_from_flags(ctrl[2], 0);
}

int spj_partition_device(const spj_tuple_t *rel, uint64_t n, int radix_bits, spj_tuple_t *out, uint32_t *pstart)
{
    if ((!rel && n) || (!out && n) || !pstart) return fail(SPJ_EINVAL, "null argument");
    if (radix_bits < 0 || radix_bits > (int)kMaxBits) return fail(SPJ_EINVAL, "radix_bits %d outside 0..%u", radix_bits, kMaxBits);
    int rc = validate_sizes(n, 0);
    if (rc) return rc;
    // an instance with an empty S whose R goes through the counted layout
    // (histogram, scan, scatter) into contiguous partitions
    InstSpec s;
    s.R = reinterpret_cast<const uint2 *>(rel);
    s.nR = (uint32_t)n;
    s.S = s.R;
    s.nS = 0;
    s.out = reinterpret_cast<uint2 *>(out);
    s.cfg = resolve(nullptr, n, n, -1.0);
    s.cfg.bits = (uint32_t)radix_bits;
    s.cfg.alg = 0;
    s.cfg.heavy = 0;
    s.cfg.cells = 0;
    s.r_contiguous = true;
    std::lock_guard<std::mutex> lk(default_context_mutex());
    Context &ctx = default_context();
    std::vector<uint32_t> counts, errs;
    hipError_t e = ctx.run({s}, counts, errs, nullptr);
    if (e == hipSuccess) e = ctx.export_partitioned_r(0, reinterpret_cast<uint2 *>(out), pstart);
    return e == hipSuccess ? SPJ_OK : hip_fail(e, "partition");
}

int spj_join_host(const spj_tuple_t *R, uint64_t nR, const spj_tuple_t *S, uint64_t nS, spj_pair_t *out,
                  uint64_t out_cap, uint64_t *out_count, const spj_config_t *cfg)
{
    if ((!R && nR) || (!S && nS) || (!out && nS) || !out_count) return fail(SPJ_EINVAL, "null argument");
    int rc = validate_sizes(nR, nS);
    if (rc) return rc;
    void *dR = nullptr, *dS = nullptr, *dO = nullptr;
    hipError_t e = hipMalloc(&dR, nR * 8 + 8);
    if (e == hipSuccess) e = hipMalloc(&dS, nS * 8 + 8);
    if (e == hipSuccess) e = hipMalloc(&dO, nS * 8 + 8);
    // page-locked staging ring on the null stream, then a wait: the join
    // runs on another stream
    if (e == hipSuccess) e = stage_h2d(dR, R, nR * 8, nullptr);
    if (e == hipSuccess) e = stage_h2d(dS, S, nS * 8, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
        rc = hip_fail(e, "spj_join_host staging");
    } else {
        uint64_t cnt = 0;
        rc = spj_join_device((const spj_tuple_t *)dR, nR, (const spj_tuple_t *)dS, nS, (spj_pair_t *)dO, nS, &cnt,
                             cfg, nullptr);
        if (rc == SPJ_OK) {
            if (cnt > out_cap) {
                rc = fail(SPJ_EOVERFLOW, "out_cap %llu < %llu pairs", (unsigned long long)out_cap, (unsigned long long)cnt);
            } else {
                e = stage_d2h(out, dO, cnt * 8, nullptr);
                if (e != hipSuccess) rc = hip_fail(e, "D2H");
                *out_count = cnt;
            }
        }
    }
    if (dR) (void)hipFree(dR);
    if (dS) (void)hipFree(dS);
    if (dO) (void)hipFree(dO);
    return rc;
}

int spj_join_batch_device(int n, const spj_tuple_t *const *R, const uint64_t *nR, const spj_tuple_t *const *S,
                          const uint64_t *nS, spj_pair_t *const *out, uint64_t *out_count, const spj_config_t *cfg,
                          spj_phase_times_t *times)
{
    if (n < 0 || (n > 0 && (!R || !nR || !S || !nS || !out || !out_count))) return fail(SPJ_EINVAL, "null argument");
    std::vector<InstSpec> specs((size_t)n);
    uint64_t total_r = 0, total_s = 0;
    for (int i = 0; i < n; i++) {   // saturating: no wrap-around
        total_r = std::min(total_r + std::min<uint64_t>(nR[i], kAutoNphjMaxBatchR + 1), kAutoNphjMaxBatchR + 1);
        total_s = std::min(total_s + std::min<uint64_t>(nS[i], uint64_t(1) << 45), uint64_t(1) << 45);
    }
    // a batch is judged by its totals; an instance with a tuner entry keeps
    // the entry's choice
    const bool nphj = auto_no_partition(cfg, total_r, total_s, kAutoNphjMaxBatchR);
    for (int i = 0; i < n; i++) {
        int rc = validate_sizes(nR[i], nS[i]);
        if (rc) return rc;
        InstSpec &s = specs[(size_t)i];
        s.R = reinterpret_cast<const uint2 *>(R[i]);
        s.S = reinterpret_cast<const uint2 *>(S[i]);
        s.out = reinterpret_cast<uint2 *>(out[i]);
        s.nR = (uint32_t)nR[i];
        s.nS = (uint32_t)nS[i];
        s.cfg = resolve(cfg, nR[i], nS[i], -1.0);
        if (nphj && !tuner_has_entry((uint32_t)nR[i], (uint32_t)nS[i], -1.0)) s.cfg.alg = SPJ_ALG_NPHJ;
        // In a batch, heavy-key handling (AUTO) from 2^14 S tuples, not 2^20:
        // it lets S take the cell layout, and a batch's histogram pass over
        // all its small S costs more than their pre-samples (4096 mixed
        // joins: 19.9 -> 19.1 ms).  A lone join of that size pays the
        // sampling and a second scatter launch for a histogram pass of a few
        // microseconds, so single calls keep 2^20.
        if ((!cfg || cfg->heavy < 0) && nS[i] >= (1u << 14) && !tuner_has_entry((uint32_t)nR[i], (uint32_t)nS[i], -1.0))
            s.cfg.heavy = 1;
        // a batch partitions in one pass (its instances' partitions larger
        // than an LDS table are joined in sub-tables)
        if (cfg && cfg->radix_passes == 2) return fail(SPJ_EINVAL, "two radix passes: single joins only");
        s.cfg.passes = 1;
        rc = check_resolved(s.cfg, nR[i]);
        if (rc) return rc;
    }
    std::lock_guard<std::mutex> lk(default_context_mutex());
    std::vector<uint32_t> counts, errs;
    PhaseTimes t;
    hipError_t e = default_context().run(specs, counts, errs, times ? &t : nullptr);
    if (e != hipSuccess) return hip_fail(e, "batch join");
    fill_times(t, times);
    int rc = SPJ_OK;
    for (int i = 0; i < n; i++) {
        out_count[i] = counts[(size_t)i];
        if (rc == SPJ_OK) rc = err_from_flags(errs[(size_t)i], (size_t)i);
    }
    return rc;
}

}  // extern "C"
