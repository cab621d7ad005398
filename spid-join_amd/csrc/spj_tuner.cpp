// spj_tuner.cpp -- the parameter table and the project-written autotuner.
//
// Instances are bucketed by (ceil log2 |R|, ceil log2 |S|/|R|, skew class);
// each bucket may hold tuned parameters (radix bits, tile size, scatter
// items per thread, LDS table size, S tuples per work item).  spj_autotune
// times candidate configurations on synthetic data of a bucket on the
// device and stores the fastest; the table can be saved to and loaded from
// a small text file.  Instances in buckets without an entry use
// default_params().
#include "spj_tuner.hpp"

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "spidjoin.h"
#include "spj_capi_internal.hpp"
#include "spj_kernels.hpp"

namespace spj {

namespace {

using Key = std::tuple<uint32_t, uint32_t, uint32_t>;

std::mutex g_mu;
std::map<Key, Resolved> g_table;

uint32_t clog2(uint64_t x)
{
    uint32_t r = 0;
    while ((uint64_t(1) << r) < x) r++;
    return r;
}

uint32_t skew_class(double s)
{
    if (s < 0) return 0;         // unknown: share the Zipf-0 bucket
    if (s < 0.75) return 0;
    if (s < 1.25) return 1;
    return 2;
}

Key bucket(uint32_t nR, uint32_t nS, double skew)
{
    uint32_t r = clog2(nR ? nR : 1);
    uint32_t ratio = nR ? clog2((nS + nR - 1) / nR) : 0;
    return Key(r, ratio, skew_class(skew));
}

}  // namespace

Resolved tuned_params(uint32_t nR, uint32_t nS, double skew_hint)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_table.find(bucket(nR, nS, skew_hint));
    if (it == g_table.end() && skew_hint >= 0) it = g_table.find(bucket(nR, nS, 0.0));
    if (it == g_table.end()) return default_params(nR, nS);
    Resolved r = it->second;
    // the heavy switch follows the instance, not the bucket representative
    Resolved d = default_params(nR, nS);
    r.heavy = d.heavy;
    return r;
}

bool tuner_has_entry(uint32_t nR, uint32_t nS, double skew_hint)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_table.count(bucket(nR, nS, skew_hint))) return true;
    return skew_hint >= 0 && g_table.count(bucket(nR, nS, 0.0));
}

}  // namespace spj

using namespace spj;

extern "C" {

/* Set the parameters of the bucket that (nR, nS, zipf) falls in. */
int spj_tuner_set(uint64_t nR, uint64_t nS, double zipf, const spj_config_t *cfg)
{
    if (!cfg) return fail(SPJ_EINVAL, "null config");
    Resolved r = default_params((uint32_t)nR, (uint32_t)nS);
    spj_config_t c = *cfg;
    Resolved x = resolve(&c, nR, nS, zipf);
    (void)r;
    int rc = check_resolved(x);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_mu);
    g_table[bucket((uint32_t)nR, (uint32_t)nS, zipf)] = x;
    return SPJ_OK;
}

/* Parameters that a single call (spj_join_device with a default config) on
 * relations of this shape would run with: the bucket's entry, else AUTO's
 * choice (the no-partition join for |R| <= 2^20) with the defaults. */
int spj_tuner_get(uint64_t nR, uint64_t nS, double zipf, spj_config_t *cfg)
{
    if (!cfg) return fail(SPJ_EINVAL, "null config");
    Resolved r = tuned_params((uint32_t)nR, (uint32_t)nS, zipf);
    if (!tuner_has_entry((uint32_t)nR, (uint32_t)nS, zipf) && auto_no_partition_single(nullptr, nR, nS))
        r.alg = SPJ_ALG_NPHJ;
    cfg->algorithm = (int32_t)r.alg;
    cfg->radix_bits = (int32_t)r.bits;
    cfg->table_log2 = (int32_t)r.table_log2;
    cfg->heavy = (int32_t)r.heavy;
    cfg->heavy_min_count = (int32_t)r.heavy_min;
    cfg->scatter_ipt = (int32_t)r.ipt;
    cfg->tile_tuples = r.tile;
    cfg->items_s_max = r.smax;
    cfg->part_layout = r.cells == 2 ? -1 : (int32_t)r.cells;
    cfg->cell_capacity_pct = r.cell_pct ? (int32_t)r.cell_pct : -1;
    cfg->radix_passes = (int32_t)r.passes;
    return SPJ_OK;
}

int spj_tuner_clear(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_table.clear();
    return SPJ_OK;
}

int spj_tuner_entries(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    return (int)g_table.size();
}

int spj_tuner_save(const char *path)
{
    FILE *f = fopen(path, "w");
    if (!f) return fail(SPJ_EINVAL, "cannot write %s", path);
    std::lock_guard<std::mutex> lk(g_mu);
    fprintf(f, "# log2R ratio_log2 skew bits table_log2 alg ipt tile smax heavy_min passes\n");
    for (auto &kv : g_table) {
        const Resolved &r = kv.second;
        fprintf(f, "%u %u %u %u %u %u %u %u %u %u %u\n", std::get<0>(kv.first), std::get<1>(kv.first),
                std::get<2>(kv.first), r.bits, r.table_log2, r.alg, r.ipt, r.tile, r.smax, r.heavy_min, r.passes);
    }
    fclose(f);
    return SPJ_OK;
}

int spj_tuner_load(const char *path)
{
    FILE *f = fopen(path, "r");
    if (!f) return fail(SPJ_EINVAL, "cannot read %s", path);
    char line[256];
    std::map<Key, Resolved> t;
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        unsigned a, b, c, bits, tl, alg, ipt, tile, smax, hm, passes = 1;   // (tables saved without passes: one)
        const int got = sscanf(line, "%u %u %u %u %u %u %u %u %u %u %u", &a, &b, &c, &bits, &tl, &alg, &ipt, &tile,
                               &smax, &hm, &passes);
        if (got != 10 && got != 11) continue;
        Resolved r;
        r.passes = passes;
        r.bits = bits;
        r.table_log2 = tl;
        r.alg = alg;
        r.ipt = ipt;
        r.tile = tile;
        r.smax = smax;
        r.heavy_min = hm;
        if (check_resolved(r) != SPJ_OK) {
            fclose(f);
            return SPJ_EINVAL;
        }
        t[Key(a, b, c)] = r;
    }
    fclose(f);
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto &kv : t) g_table[kv.first] = kv.second;
    return SPJ_OK;
}

/* Tune the bucket of (R, S, zipf) on the given device-resident relations and
 * keep the fastest configuration.  Coordinate descent from the defaults: one
 * parameter at a time takes each of its candidate values with the others at
 * their best so far, in the order
 *   algorithm      partitioned hash join, or the no-partition join (|R| <= 2^24)
 *   radix bits     default - 2 .. default + 1 (R partitions within 4 LDS tables)
 *   scatter ipt    8, 10, 12 tuples per thread per scatter round
 *   table buckets  2^12, 2^13
 *   S per item     2^16, 2^17, 2^18 S tuples per join work item
 *   passes         1 or 2 radix passes, where one pass of the most bits
 *                  leaves R partitions larger than an LDS table
 * and the sweep runs twice.  Each candidate runs reps + 1 times (the first
 * is a warm-up); its score is the median device time (events around the
 * launch sequence).  Workgroup sizes are not searched: the scatter and join
 * kernels run one 1024-thread workgroup per CU that holds the whole LDS.
 * best_ms receives the winner's median time. */
int spj_autotune(const spj_tuple_t *R, uint64_t nR, const spj_tuple_t *S, uint64_t nS, spj_pair_t *out, double zipf,
                 int reps, float *best_ms)
{
    if (!R || !S || !out || nR == 0 || nS == 0) return fail(SPJ_EINVAL, "bad autotune arguments");
    int rc = validate_sizes(nR, nS);
    if (rc) return rc;
    if (reps < 1) reps = 1;
    const Resolved base = default_params((uint32_t)nR, (uint32_t)nS);
    std::lock_guard<std::mutex> lk(default_context_mutex());
    std::map<std::vector<uint32_t>, float> seen;   // candidate -> median ms
    auto score = [&](const Resolved &c, float *ms) -> int {
        const std::vector<uint32_t> key = {c.alg, c.bits, c.ipt, c.table_log2, c.smax, c.passes};
        auto it = seen.find(key);
        if (it != seen.end()) {
            *ms = it->second;
            return SPJ_OK;
        }
        InstSpec s;
        s.R = reinterpret_cast<const uint2 *>(R);
        s.S = reinterpret_cast<const uint2 *>(S);
        s.out = reinterpret_cast<uint2 *>(out);
        s.nR = (uint32_t)nR;
        s.nS = (uint32_t)nS;
        s.cfg = c;
        std::vector<float> ts;
        for (int i = 0; i < reps + 1; i++) {
            PhaseTimes t;
            uint32_t count = 0, err = 0;
            hipError_t e = run_one(s, count, err, &t);
            if (e != hipSuccess) return hip_fail(e, "autotune run");
            if (err) break;   // e.g. a duplicate key: no score
            if (i) ts.push_back(t.total);
        }
        float med = 1e30f;
        if (!ts.empty()) {
            std::sort(ts.begin(), ts.end());
            med = ts[ts.size() / 2];
        }
        seen[key] = med;
        *ms = med;
        return SPJ_OK;
    };
    Resolved win = base;
    float best = 1e30f;
    rc = score(win, &best);
    if (rc) return rc;
    auto try_dim = [&](auto set, std::initializer_list<uint32_t> vals) -> int {
        for (uint32_t v : vals) {
            Resolved c = win;
            set(c, v);
            if (c.alg == SPJ_ALG_PHJ && c.passes == 1 && (nR >> c.bits) > 4ull * join_table_cap()) continue;
            if (c.passes == 2 && (c.alg == SPJ_ALG_NPHJ || c.bits == 0)) continue;
            if (c.alg == SPJ_ALG_NPHJ && nR > (uint64_t(1) << 24)) continue;
            float ms = 0;
            int r = score(c, &ms);
            if (r) return r;
            if (ms < best) {
                best = ms;
                win = c;
            }
        }
        return SPJ_OK;
    };
    for (int sweep = 0; sweep < 2; sweep++) {
        if ((rc = try_dim([](Resolved &c, uint32_t v) { c.alg = v; }, {(uint32_t)SPJ_ALG_PHJ, (uint32_t)SPJ_ALG_NPHJ})))
            return rc;
        if (win.alg == SPJ_ALG_NPHJ) break;   // no partition parameters to tune
        const uint32_t b0 = base.bits;
        std::vector<uint32_t> bits;
        for (int d = -2; d <= 1; d++)
            if ((int)b0 + d >= 0 && (int)b0 + d <= (int)kMaxBits) bits.push_back(b0 + d);
        for (uint32_t b : bits)
            if ((rc = try_dim([](Resolved &c, uint32_t v) { c.bits = v; }, {b}))) return rc;
        if ((rc = try_dim([](Resolved &c, uint32_t v) { c.ipt = v; }, {8u, 10u, 12u}))) return rc;
        if ((rc = try_dim([](Resolved &c, uint32_t v) { c.table_log2 = v; }, {12u, 13u}))) return rc;
        if ((rc = try_dim([](Resolved &c, uint32_t v) { c.smax = v; }, {1u << 16, 1u << 17, 1u << 18}))) return rc;
        if ((nR >> kMaxBits) > join_table_cap())
            if ((rc = try_dim([](Resolved &c, uint32_t v) { c.passes = v; }, {1u, 2u}))) return rc;
    }
    {
        std::lock_guard<std::mutex> lk2(g_mu);
        g_table[bucket((uint32_t)nR, (uint32_t)nS, zipf)] = win;
    }
    if (best_ms) *best_ms = best;
    return SPJ_OK;
}

}  // extern "C"
