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

/* Parameters that an instance of this shape would run with. */
int spj_tuner_get(uint64_t nR, uint64_t nS, double zipf, spj_config_t *cfg)
{
    if (!cfg) return fail(SPJ_EINVAL, "null config");
    Resolved r = tuned_params((uint32_t)nR, (uint32_t)nS, zipf);
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
    fprintf(f, "# log2R ratio_log2 skew bits table_log2 alg ipt tile smax heavy_min\n");
    for (auto &kv : g_table) {
        const Resolved &r = kv.second;
        fprintf(f, "%u %u %u %u %u %u %u %u %u %u\n", std::get<0>(kv.first), std::get<1>(kv.first),
                std::get<2>(kv.first), r.bits, r.table_log2, r.alg, r.ipt, r.tile, r.smax, r.heavy_min);
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
        unsigned a, b, c, bits, tl, alg, ipt, tile, smax, hm;
        if (sscanf(line, "%u %u %u %u %u %u %u %u %u %u", &a, &b, &c, &bits, &tl, &alg, &ipt, &tile, &smax, &hm) != 10)
            continue;
        Resolved r;
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

/* Time candidate configurations for the bucket of (R, S, zipf) on the given
 * device-resident sample relations and keep the fastest.  Candidates vary
 * radix bits around the default (-2..+1), scatter tuples per thread (8, 12)
 * and LDS table bucket count (2^12, 2^13).  reps timed runs per candidate (median).
 * best_ms receives the winner's median time. */
int spj_autotune(const spj_tuple_t *R, uint64_t nR, const spj_tuple_t *S, uint64_t nS, spj_pair_t *out, double zipf,
                 int reps, float *best_ms)
{
    if (!R || !S || !out || nR == 0 || nS == 0) return fail(SPJ_EINVAL, "bad autotune arguments");
    if (reps < 1) reps = 1;
    Resolved base = default_params((uint32_t)nR, (uint32_t)nS);
    std::vector<Resolved> cands;
    for (int db = -2; db <= 1; db++) {
        int b = (int)base.bits + db;
        if (b < 0 || b > (int)kMaxBits) continue;
        for (uint32_t ipt : {8u, 12u}) {
            for (uint32_t tl : {12u, 13u}) {
                Resolved r = base;
                r.bits = (uint32_t)b;
                r.ipt = ipt;
                r.table_log2 = tl;
                // R partitions must fit the LDS table in at most a few sub-tables
                uint64_t per = nR >> r.bits;
                if (per > 4ull * join_table_cap()) continue;
                cands.push_back(r);
            }
        }
    }
    float best = 1e30f;
    Resolved win = base;
    std::lock_guard<std::mutex> lk(default_context_mutex());
    for (const Resolved &c : cands) {
        InstSpec s;
        s.R = reinterpret_cast<const uint2 *>(R);
        s.S = reinterpret_cast<const uint2 *>(S);
        s.out = reinterpret_cast<uint2 *>(out);
        s.nR = (uint32_t)nR;
        s.nS = (uint32_t)nS;
        s.cfg = c;
        s.cfg.heavy = default_params((uint32_t)nR, (uint32_t)nS).heavy;
        std::vector<float> ts;
        std::vector<uint32_t> counts, errs;
        for (int i = 0; i < reps + 1; i++) {
            PhaseTimes t;
            hipError_t e = default_context().run({s}, counts, errs, &t);
            if (e != hipSuccess) return hip_fail(e, "autotune run");
            if (errs[0]) break;
            if (i) ts.push_back(t.total);   // first run is a warm-up
        }
        if (ts.empty()) continue;
        std::sort(ts.begin(), ts.end());
        float med = ts[ts.size() / 2];
        if (med < best) {
            best = med;
            win = c;
        }
    }
    {
        std::lock_guard<std::mutex> lk2(g_mu);
        g_table[bucket((uint32_t)nR, (uint32_t)nS, zipf)] = win;
    }
    if (best_ms) *best_ms = best;
    return SPJ_OK;
}

}  // extern "C"
