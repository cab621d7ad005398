// spj_engine.cpp -- batch planning and phase launch sequence.
//
// Phase order follows the reference's plans (execution.hpp:15-157 and the
// compound operators of queryCompoundFunc.cc): partition R, partition S
// (with SPID's skew replication folded in as heavy-key handling), then the
// per-partition join.  Rank threads, barriers and the host shuffle of the
// reference (pipeline.cc, rotate_and_stream.c) have no counterpart: every
// phase is a kernel on one stream and partitions are written in place in HBM.
#include "spj_engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "spj_kernels.hpp"

namespace spj {

#define SPJ_TRY(x)                                  \
    do {                                            \
        hipError_t _e = (x);                        \
        if (_e != hipSuccess) return _e;            \
    } while (0)

hipError_t PinBuf::ensure(size_t bytes)
{
    if (bytes <= cap && p) return hipSuccess;
    if (p) {
        hipError_t e = hipHostFree(p);
        p = nullptr;
        cap = 0;
        if (e != hipSuccess) return e;
    }
    size_t want = std::max<size_t>(bytes, 4096);
    want = (want + 65535) & ~size_t(65535);
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        p = nullptr;
        return e;
    }
    cap = want;
    return hipSuccess;
}

void PinBuf::release()
{
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
}

hipError_t DevBuf::ensure(size_t bytes)
{
    if (bytes <= cap && p) return hipSuccess;
    if (p) {
        hipError_t e = hipFree(p);
        p = nullptr;
        cap = 0;
        if (e != hipSuccess) return e;
    }
    size_t want = std::max<size_t>(bytes, 256);
    want = (want + 4095) & ~size_t(4095);
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        return e;
    }
    cap = want;
    return hipSuccess;
}

void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
}

const DeviceInfo &device_info()
{
    static DeviceInfo info;
    static std::once_flag once;
    std::call_once(once, [] {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return;
        info.device = dev;
        info.cus = prop.multiProcessorCount;
        info.lds_per_block = prop.sharedMemPerBlock;
        info.arch = prop.gcnArchName;
    });
    return info;
}

static uint32_t ceil_log2(uint64_t x)
{
    uint32_t r = 0;
    while ((uint64_t(1) << r) < x) r++;
    return r;
}

// Tiles of at least one scatter round, and about 2048 of them per relation
// at most: enough for every scatter workgroup to take several, few enough
// that the histogram's counter atomics stay cheap.
uint32_t default_tile(uint64_t n)
{
    uint64_t t = 8192;
    while (n / t > 2048) t <<= 1;
    return (uint32_t)t;
}

Resolved default_params(uint32_t nR, uint32_t nS)
{
    Resolved r;
    r.table_log2 = 13;   // 8192 buckets: about 2 R tuples per bucket at 16K tuples
    // Aim at <= 16K R tuples per partition: one LDS table (17408 tuples) with
    // room for the binomial spread of partition sizes.
    uint32_t b = nR > 16384 ? ceil_log2((uint64_t(nR) + 16383) / 16384) : 0;
    r.bits = std::min<uint32_t>(b, kMaxBits);
    r.ipt = 12;
    r.tile = default_tile(std::max(nR, nS));
    r.smax = 131072;   // at 16K R tuples per partition: several S tuples per table build
    r.heavy = nS >= (1u << 20) ? 1u : 0u;   // (a batch lowers this to 2^14; see spj_join_batch_device)
    r.heavy_min = 16;
    r.alg = 0;
    // Two passes when one pass of kMaxBits would leave R partitions of four
    // or more LDS tables: the join would read each partition's S that many
    // times, against one more read and write of R and S for a second pass.
    const uint64_t per = (uint64_t(nR) + (1u << kMaxBits) - 1) >> kMaxBits;
    r.passes = per > 3ull * 16384 ? 2u : 1u;
    return r;
}

Context::Context()
{
    const char *g = getenv("SPJ_GRAPHS");   // SPJ_GRAPHS=0: never replay launch sequences as graphs
    graphs_ = !(g && g[0] == '0');
    (void)hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking);
    // phase-timing events record without a system-scope fence: a fenced
    // record writes back and invalidates the caches (several microseconds
    // each, and the next kernel starts cold); only the join's end needs to
    // be visible to the host, and the result copy's synchronisation does that
    for (auto &e : ev_) (void)hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
    (void)hipEventCreateWithFlags(&desc_ev_, hipEventDisableTiming);
    // page-locked staging is slow to allocate: take a size that covers
    // most batches now, not in the first join (a latency outlier)
    (void)pin_desc_.ensure(size_t(1) << 20);
    (void)pin_ctrl_.ensure(size_t(64) << 10);
}

Context::~Context()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graph();
    for (auto &e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (desc_ev_) (void)hipEventDestroy(desc_ev_);
    DevBuf *bufs[] = {&desc_, &counts_, &pstart_, &rp_, &sp_, &items_, &ctrl_,
                      &heavy_, &cand_, &holes_, &tmp_, &segbuf_, &tiles_, &nphj_, &ext_};
    for (auto *b : bufs) b->release();
    pin_desc_.release();
    pin_ctrl_.release();
    if (stream_) (void)hipStreamDestroy(stream_);
}

// ctrl layout (uint32): [0] nitems, [1] qhead, [2] holes recorded, [3] pad,
// [4, 20) the tile queues of the two scatter launches, then n words each of: out_lo (pairs reserved by join items, from the front),
// out_hi (heavy pairs, from the back), out_fill (pairs written by items),
// err, nheavy, spart and rpart (S and R tuples in partitions).
enum {
    kCtrlLo = 0, kCtrlHi = 1, kCtrlFill = 2, kCtrlErr = 3, kCtrlHeavy = 4, kCtrlSPart = 5, kCtrlRPart = 6,
    kCtrlArrays = 7
};
constexpr size_t kCtrlHead = 4 + 2 * kScatterGroups;
static inline size_t ctrl_words(size_t n) { return kCtrlHead + kCtrlArrays * n; }
static inline size_t ctrl_at(size_t n, int which, size_t i) { return kCtrlHead + (size_t)which * n + i; }

// Capacity of one S cell that takes a group's tuples of one digit, from the
// group's tuple count n and the fan-out F: by default 1.5 (x2: 2) times the
// mean fill plus 8 standard deviations of a uniform spread plus 64, so that
// only keys too frequent for the heavy detection to have missed overflow a
// cell.
static uint32_t cell_capacity(uint64_t n, uint32_t F, uint32_t pct, bool x2)
{
    if (!n) return 0;
    const double m = (double)n / F;
    if (pct) return (uint32_t)std::ceil(m * pct / 100.0) + 1u;
    return (uint32_t)std::ceil((x2 ? 2.0 : 1.5) * m + 8.0 * std::sqrt(m) + 64.0);
}

static bool same_params(const Resolved &a, const Resolved &b)
{
    return a.bits == b.bits && a.table_log2 == b.table_log2 && a.alg == b.alg && a.heavy == b.heavy &&
           a.heavy_min == b.heavy_min && a.ipt == b.ipt && a.tile == b.tile && a.smax == b.smax && a.cells == b.cells &&
           a.cell_pct == b.cell_pct && a.cell_mean_x2 == b.cell_mean_x2 && a.passes == b.passes;
}

static bool same_specs(const std::vector<InstSpec> &a, const std::vector<InstSpec> &b)
{
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); i++) {
        const InstSpec &x = a[i], &y = b[i];
        if (x.R != y.R || x.nR != y.nR || x.S != y.S || x.nS != y.nS || x.out != y.out || x.Rp_pre != y.Rp_pre ||
            x.pst_pre != y.pst_pre || x.r_contiguous != y.r_contiguous || x.s_contiguous != y.s_contiguous || x.shift != y.shift ||
            x.partition_only != y.partition_only || !same_params(x.cfg, y.cfg) || x.rseg.size() != y.rseg.size() ||
            x.sseg.size() != y.sseg.size())
            return false;
        for (size_t k = 0; k < x.rseg.size(); k++)
            if (x.rseg[k].x != y.rseg[k].x || x.rseg[k].y != y.rseg[k].y) return false;
        for (size_t k = 0; k < x.sseg.size(); k++)
            if (x.sseg[k].x != y.sseg[k].x || x.sseg[k].y != y.sseg[k].y) return false;
    }
    return true;
}

void Context::drop_graph()
{
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    graph_exec_ = nullptr;
}

// A call that repeats the previous call's specs exactly (same pointers,
// sizes and parameters) replays the launch sequence as a HIP graph: the
// planning, the descriptor contents and every buffer address are the same,
// so only the host's launch work is saved.  The graph is captured at the
// first repeat and dropped by any other call (which may move buffers).  The
// graph holds no descriptor upload: the previous call, identical and
// launched directly, left the same descriptors on the device, and no
// kernel writes them.  Timed calls (phase events) and the counted redo
// launch directly: event records captured in a graph kept the times of
// its first replay on MI355X (ROCm 7), every later replay reported them
// again.
hipError_t Context::enqueue(const std::vector<InstSpec> &specs_in, PhaseTimes *timed)
{
    if (graphs_ && !timed && !force_counted_ && last_valid_ && &specs_in != &h_specs_ && same_specs(specs_in, h_specs_)) {
        if (!graph_exec_) {
            hipError_t e = hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed);
            if (e == hipSuccess) {
                capturing_ = true;
                e = enqueue_plan(h_specs_, timed);
                capturing_ = false;
                hipGraph_t g = nullptr;
                const hipError_t e2 = hipStreamEndCapture(stream_, &g);
                if (e == hipSuccess) e = e2;
                if (e == hipSuccess) e = hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0);
                if (g) (void)hipGraphDestroy(g);
                if (e != hipSuccess) graph_exec_ = nullptr;
                else graph_captures_++;
            }
            if (e != hipSuccess) {   // no graph for this call: launch directly, and stop trying
                (void)hipGetLastError();
                graphs_ = false;
                last_valid_ = false;
                return enqueue(specs_in, timed);
            }
        }
        SPJ_TRY(hipGraphLaunch(graph_exec_, stream_));
        graph_replays_++;
        return hipSuccess;
    }
    drop_graph();
    last_valid_ = false;
    const hipError_t e = enqueue_plan(specs_in, timed);
    last_valid_ = e == hipSuccess && !force_counted_;
    return e;
}

hipError_t Context::enqueue_plan(const std::vector<InstSpec> &specs_in, PhaseTimes *timed)
{
    if (&specs_in != &h_specs_) h_specs_ = specs_in;
    const std::vector<InstSpec> &specs = h_specs_;
    const size_t n = specs.size();
    h_insts_.assign(n, InstDesc{});
    h_cells_.clear();
    h_segs_.clear();
    max_bits_ = 0;
    max_table_ = 0;
    max_ipt_ = 4;
    any_heavy_ = false;

    // ---- plan: regions, tiles, scan segments
    // Tiles are listed as [R cells | R counted | S counted | S cells]: the
    // histogram pass reads the counted ones (a contiguous range); when R
    // has cells and heavy keys, R is scattered by a launch of its own (its
    // tiles record the heavy keys' R rids that the S scatter emits).
    // counters: the counted relations' first (zeroed), then the cell
    // relations' (set to the cell bases), so the two never overlap
    uint64_t cnt_counted = 0, cnt_cells = 0;
    uint64_t rp_total = 0, sp_total = 0, pst_total = 0, items_bound = 0;
    uint32_t n_joined = 0;   // instances joined through partitions (not by the no-partition join)
    // One S-range size per batch (the item kernel takes a single value):
    // the configured one, but small enough for about two items per CU when
    // the batch's S is small (each item builds its R table again, but the
    // items run side by side: a lone 2^16 x 2^18 join has four partitions,
    // and four items left k_join at one item's time on one CU)
    uint32_t smax = specs.empty() ? 131072u : specs[0].cfg.smax;
    {
        uint64_t stot = 0;
        for (const InstSpec &s : specs) stot += s.nS;
        const uint64_t per = (stot + 2ull * device_info().cus - 1) / (2ull * std::max(1, device_info().cus));
        smax = (uint32_t)std::min<uint64_t>(smax, std::max<uint64_t>(per, 4096));
    }
    uint32_t nheavy_inst = 0;
    std::vector<uint64_t> rp_off(n), sp_off(n);
    std::vector<int> heavy_idx(n, -1);
    std::vector<char> cells[2] = {std::vector<char>(n, 0), std::vector<char>(n, 0)};
    std::vector<uint64_t> cnt0[2] = {std::vector<uint64_t>(n, 0), std::vector<uint64_t>(n, 0)};
    std::vector<TileRun> lists[4];   // R cells, R counted, S counted, S cells
    bool split = false;
    for (size_t i = 0; i < n; i++) {
        const InstSpec &s = specs[i];
        const Resolved &c = s.cfg;
        const uint32_t F = 1u << c.bits;
        max_bits_ = std::max(max_bits_, c.bits);
        max_table_ = std::max(max_table_, c.table_log2);
        max_ipt_ = std::max<int>(max_ipt_, (int)c.ipt);
        const bool pre = s.Rp_pre != nullptr;
        // no-partition hash join: no heavy keys, no partition tiles (its
        // partitions stay empty, so it has no work items either)
        const bool nphj = c.alg == 1 && !pre;
        // cells: no histogram pass.  Cursors are 32-bit, so a relation's
        // cell region plus all its tuples stays below 2^32.  Automatically
        // for large hash-join relations; for S only with heavy-key handling,
        // which takes the keys frequent enough to overflow a cell (R keys
        // are unique).
        const bool segmented = !s.rseg.empty() || !s.sseg.empty();
        if (segmented && (pre || nphj)) return hipErrorInvalidValue;
        const bool heavy_on = c.heavy && s.nS > 0 && !pre && !nphj && !segmented;
        // AUTO: cells for R of any size (unique keys never crowd a cell; a
        // batch of 4096 mixed joins, all R below 2^20: 21.2 -> 20.1 ms
        // without their histogram pass), for S from 2^14 tuples and only
        // with heavy-key handling, whose pre-sample takes the keys that
        // could overflow a cell (one skewed S overflowing a cell would redo
        // its whole batch counted).  AUTO turns heavy handling on from 2^20
        // S tuples in single calls, from 2^14 in batches.
        auto use_cells = [&](uint32_t len, bool skew_safe, uint32_t min_len) {
            return !force_counted_ && len > 0 && len <= (1u << 30) && c.cell_pct <= 200 &&
                   (c.cells == 1 || (c.cells == 2 && (c.alg == 0 || c.alg == 2) && len >= min_len && skew_safe));
        };
        // a first pass of two may leave its partitions in cells too: the
        // second pass's instances read them as segments
        if (s.partition_only != specs[0].partition_only || (s.partition_only && (nphj || pre)))
            return hipErrorInvalidValue;
        if (s.shift + c.bits > kMaxDigitBits) return hipErrorInvalidValue;
        cells[0][i] = !pre && !nphj && !s.r_contiguous && use_cells(s.nR, true, 1u);
        cells[1][i] = !nphj && !s.s_contiguous && use_cells(s.nS, heavy_on, 1u << 14);
        InstDesc &d = h_insts_[i];
        d.R = s.R;
        d.S = s.S;
        d.out = s.out;
        d.nR = s.nR;
        d.nS = s.nS;
        d.bits = c.bits;
        d.shift = s.shift;
        d.heavy = heavy_on ? 1u : 0u;
        d.table_log2 = c.table_log2;
        d.alg = (c.alg == 1 && !nphj) ? 0u : c.alg;   // a resident table is partitioned: hash join
        d.pstartR = pst_total;
        d.pstartS = pst_total + F + 1;
        pst_total += 2ull * (F + 1);
        if (d.heavy) {
            heavy_idx[i] = (int)nheavy_inst++;
            any_heavy_ = true;
            if (cells[0][i]) split = true;
        }
        items_bound += F + (uint64_t)s.nS / smax + 1;
        if (!nphj) n_joined++;
        for (uint32_t rel = pre ? 1u : 0u; rel < 2; rel++) {
            const uint32_t len = nphj ? 0u : (rel ? s.nS : s.nR);
            const bool cl = cells[rel][i];
            // cells: two scatter rounds per tile, handed out by the scatter's
            // tile queues in small steps so the workgroups finish together;
            // counted: larger tiles, each costing F counter atomics in the
            // histogram pass
            const uint32_t tile = cl ? 2u * kScatterNT * c.ipt : c.tile;
            uint32_t T = (len + tile - 1) / tile;
            if (!(rel ? s.sseg : s.rseg).empty()) {
                uint64_t tot = 0;
                T = 0;
                for (const uint2 &sg : rel ? s.sseg : s.rseg) {
                    T += (sg.y + tile - 1) / tile;
                    tot += sg.y;
                }
                if (tot != len) return hipErrorInvalidValue;   // the segments make up the relation
            }
            uint64_t &cnt_total = cl ? cnt_cells : cnt_counted;   // cell offsets are shifted below
            cnt0[rel][i] = cnt_total;
            if (!cl) {
                ScanSeg sg{};
                sg.inst = (uint32_t)i;
                sg.rel = rel;
                sg.bits = c.bits;
                sg.t_count = kScatterGroups;
                sg.cnt_off0 = cnt_total;
                sg.pstart_off = rel ? d.pstartS : d.pstartR;
                h_segs_.push_back(sg);
            }
            if (T) {
                TileRun ru{};
                ru.inst = (uint32_t)i;
                ru.rel = rel;
                ru.ntiles = T;
                ru.tile = tile;
                ru.len = len;
                ru.cnt_off0 = cnt_total;
                ru.bits = c.bits;
                auto &lst = lists[rel == 0 ? (cl ? 0 : 1) : (cl ? 3 : 2)];
                const std::vector<uint2> &segv = rel ? s.sseg : s.rseg;
                if (segv.empty()) {
                    lst.push_back(ru);
                } else {   // one run per segment, sharing the relation's counters and cells
                    for (const uint2 &sg : segv) {
                        if (!sg.y) continue;
                        ru.base = sg.x;
                        ru.len = sg.y;
                        ru.ntiles = (sg.y + tile - 1) / tile;
                        lst.push_back(ru);
                    }
                }
            }
            cnt_total += (uint64_t)kScatterGroups * F;
        }
    }
    const uint64_t cnt_total = cnt_counted + cnt_cells;
    for (size_t i = 0; i < n; i++)
        for (uint32_t rel = 0; rel < 2; rel++)
            if (cells[rel][i]) cnt0[rel][i] += cnt_counted;
    for (int l : {0, 3})
        for (TileRun &ru : lists[l]) ru.cnt_off0 += cnt_counted;
    // tile-list positions
    uint64_t pos = 0, lsize[4] = {0, 0, 0, 0};
    h_runs_.clear();
    for (int l = 0; l < 4; l++)
        for (TileRun &ru : lists[l]) {
            ru.first = (uint32_t)pos;
            pos += ru.ntiles;
            lsize[l] += ru.ntiles;
        }
    if (pos > 0x7FFFFFFFull) return hipErrorInvalidValue;
    ntiles_ = (uint32_t)pos;
    const uint32_t hist_first = (uint32_t)lsize[0];
    const uint32_t n_counted = (uint32_t)(lsize[1] + lsize[2]);
    // scatter launches: [0, scat_split) and [scat_split, ntiles_)
    const uint32_t scat_split = split ? (uint32_t)(lsize[0] + lsize[1]) : 0u;
    for (auto &l : lists)
        for (TileRun &ru : l) {
            const bool second = ru.first >= scat_split;
            ru.launch0 = second ? scat_split : 0u;
            ru.launch_n = second ? ntiles_ - scat_split : scat_split;
            h_runs_.push_back(ru);
        }
    // regions: contiguous, or cells sized by each group's share of the
    // relation's tuples
    auto gk = [](size_t i, uint32_t rel, uint32_t g) { return (i * 2 + rel) * kScatterGroups + g; };
    std::vector<uint64_t> gsize((size_t)n * 2 * kScatterGroups, 0);
    for (const TileRun &ru : h_runs_) {
        if (!cells[ru.rel][ru.inst]) continue;
        const uint64_t f = ru.first - ru.launch0, e = f + ru.ntiles;   // positions in the launch
        for (uint32_t g = 0; g < kScatterGroups; g++) {
            // positions p in [f, e) with tile_group(p) == g
            auto below = [g](uint64_t x) { return (x + kScatterGroups - 1 - g) / kScatterGroups; };
            uint64_t tuples = (below(e) - below(f)) * ru.tile;
            if (tile_group((uint32_t)(e - 1)) == g) tuples -= (uint64_t)ru.ntiles * ru.tile - ru.len;   // short last tile
            gsize[gk(ru.inst, ru.rel, g)] += tuples;
        }
    }
    std::vector<uint32_t> gcell0(gsize.size(), 0), gcap(gsize.size(), 0);
    for (size_t i = 0; i < n; i++) {
        InstDesc &d = h_insts_[i];
        const uint32_t F = 1u << d.bits;
        for (uint32_t rel = 0; rel < 2; rel++) {
            uint64_t &total = rel ? sp_total : rp_total;
            (rel ? sp_off : rp_off)[i] = total;
            if (!cells[rel][i]) {
                if (d.alg != 1 && (rel == 1 || !specs[i].Rp_pre)) total += rel ? d.nS : d.nR;
                continue;
            }
            uint64_t region = 0;
            const uint32_t first = (uint32_t)h_cells_.size() + 1;
            for (uint32_t g = 0; g < kScatterGroups; g++) {
                const size_t k = gk(i, rel, g);
                gcap[k] = cell_capacity(gsize[k], F, specs[i].cfg.cell_pct, specs[i].cfg.cell_mean_x2 != 0);
                gcell0[k] = (uint32_t)region;
                region += (uint64_t)F * gcap[k];
                CellGroup cg{};
                cg.cnt_off = cnt0[rel][i] + (uint64_t)g * F;
                cg.cell0 = gcell0[k];
                cg.cap = gcap[k];
                cg.bits = d.bits;
                h_cells_.push_back(cg);
            }
            (rel ? d.s_cells : d.r_cells) = first;
            (rel ? d.s_sink : d.r_sink) = (uint32_t)region;
            total += region + kSinkTuples;
        }
    }
    for (TileRun &ru : h_runs_)
        if (cells[ru.rel][ru.inst]) ru.cells = (ru.rel ? h_insts_[ru.inst].s_cells : h_insts_[ru.inst].r_cells);

    // ---- heavy-key jobs (sampling blocks of both phases)
    h_jobs_.clear();
    uint32_t hblocks[2] = {0, 0};
    for (size_t i = 0; i < n; i++) {
        if (heavy_idx[i] < 0) continue;
        HeavyJob j{};
        j.inst = (uint32_t)i;
        j.nblocks = heavy_sample_blocks(specs[i].nS, 1024, &j.stride);
        const uint32_t pre = (j.nblocks + kPreSample - 1) / kPreSample;
        j.first_block[0] = hblocks[0];
        j.first_block[1] = hblocks[1];
        j.min_count = specs[i].cfg.heavy_min;
        j.cand_off = (uint64_t)heavy_idx[i] * cand_words();
        hblocks[0] += pre;
        // The second sample only where the pre-sample can miss a key that
        // overflows a cell.  Such a key holds more than half a cell's mean
        // fill in each of the 8 groups: a share of S above 1 / (2 F).  The
        // pre-sample (kLocalMin in an 8192-key block) sees shares from
        // about 1 / 2048 on, so with F < 1024 partitions it finds every one;
        // the second sample then only adds milder keys, whose direct
        // emission did not pay for the sampling (4096 mixed joins, |R| <=
        // 2^20: 20.4 -> 19.5 ms without it).  At 2^26 x 2^28, Zipf 1.0
        // (F = 4096) the pre-sample alone left a key that overflowed a cell
        // (5.2 ms with the counted redo, 2.36 ms with the second sample).
        if (specs[i].cfg.bits >= 10) hblocks[1] += j.nblocks - pre;
        h_jobs_.push_back(j);
    }

    // ---- no-partition hash joins: one table of >= 2|R| slots each
    h_nphj_.clear();
    uint64_t nphj_words = 0, nphj_bb = 0, nphj_pb = 0;
    for (size_t i = 0; i < n; i++) {
        if (h_insts_[i].alg != 1 || specs[i].Rp_pre) continue;
        NphjJob j{};
        j.inst = (uint32_t)i;
        j.log2cap = 10;
        while ((1ull << j.log2cap) < 2ull * specs[i].nR) j.log2cap++;
        if (j.log2cap > 31) return hipErrorInvalidValue;   // the table's slot index is 32-bit: |R| <= 2^30
        j.table = nphj_words;
        j.first_build = (uint32_t)nphj_bb;
        j.first_probe = (uint32_t)nphj_pb;
        nphj_words += 1ull << j.log2cap;
        nphj_bb += (specs[i].nR + kNphjBuildTuples - 1) / kNphjBuildTuples;
        nphj_pb += (specs[i].nS + kNphjProbeTuples - 1) / kNphjProbeTuples;
        h_nphj_.push_back(j);
    }
    if (nphj_bb > 0x7FFFFFFFull || nphj_pb > 0x7FFFFFFFull) return hipErrorInvalidValue;

    // ---- buffers
    // descriptor block layout (256-byte aligned arrays)
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t o_insts = 0;
    const size_t o_runs = o_insts + al(n * sizeof(InstDesc));
    const size_t o_segs = o_runs + al(h_runs_.size() * sizeof(TileRun));
    const size_t o_jobs = o_segs + al(h_segs_.size() * sizeof(ScanSeg));
    const size_t o_cells = o_jobs + al(h_jobs_.size() * sizeof(HeavyJob));
    const size_t o_nphj = o_cells + al(h_cells_.size() * sizeof(CellGroup));
    const size_t desc_bytes = o_nphj + al(h_nphj_.size() * sizeof(NphjJob));
    SPJ_TRY(desc_.ensure(desc_bytes));
    SPJ_TRY(counts_.ensure(cnt_total * 4));
    SPJ_TRY(pstart_.ensure(pst_total * 4));
    SPJ_TRY(rp_.ensure(rp_total * 8));
    SPJ_TRY(sp_.ensure(sp_total * 8));
    SPJ_TRY(items_.ensure(items_bound * sizeof(WorkItem)));
    SPJ_TRY(tiles_.ensure((size_t)ntiles_ * sizeof(TileDesc)));
    SPJ_TRY(holes_.ensure(items_bound * sizeof(uint4)));
    max_holes_ = (uint32_t)std::min<uint64_t>(items_bound, 0xFFFFFFFFull);
    SPJ_TRY(ctrl_.ensure(ctrl_words(n) * 4));
    if (nphj_words) SPJ_TRY(nphj_.ensure(nphj_words * 8));
    if (nheavy_inst) {
        SPJ_TRY(heavy_.ensure((size_t)nheavy_inst * kHeavyWords * 4));
        SPJ_TRY(cand_.ensure((size_t)nheavy_inst * cand_words() * 4));
    }
    for (size_t i = 0; i < n; i++) {
        InstDesc &d = h_insts_[i];
        d.Rp = specs[i].Rp_pre ? const_cast<uint2 *>(specs[i].Rp_pre) : rp_.as<uint2>() + rp_off[i];
        d.Sp = sp_.as<uint2>() + sp_off[i];
        d.heavy_count = ctrl_.as<uint32_t>() + ctrl_at(n, kCtrlHeavy, i);
        if (heavy_idx[i] >= 0) {
            d.heavy_keys = heavy_.as<uint32_t>() + (size_t)heavy_idx[i] * kHeavyWords;
            d.heavy_rids = d.heavy_keys + kHeavySlots;
        }
    }

    hipStream_t st = stream_;
    // one upload of all descriptors through page-locked staging; the
    // staging is reused once the previous upload has left it.  A capture
    // uploads nothing: the identical previous call left the descriptors on
    // the device (see enqueue).
    if (!capturing_) {
        if (desc_pending_) SPJ_TRY(hipEventSynchronize(desc_ev_));
        SPJ_TRY(pin_desc_.ensure(desc_bytes));
        {
            char *h = pin_desc_.as<char>();
            if (n) memcpy(h + o_insts, h_insts_.data(), n * sizeof(InstDesc));
            if (!h_runs_.empty()) memcpy(h + o_runs, h_runs_.data(), h_runs_.size() * sizeof(TileRun));
            if (!h_segs_.empty()) memcpy(h + o_segs, h_segs_.data(), h_segs_.size() * sizeof(ScanSeg));
            if (!h_jobs_.empty()) memcpy(h + o_jobs, h_jobs_.data(), h_jobs_.size() * sizeof(HeavyJob));
            if (!h_cells_.empty()) memcpy(h + o_cells, h_cells_.data(), h_cells_.size() * sizeof(CellGroup));
            if (!h_nphj_.empty()) memcpy(h + o_nphj, h_nphj_.data(), h_nphj_.size() * sizeof(NphjJob));
        }
        SPJ_TRY(hipMemcpyAsync(desc_.p, pin_desc_.p, desc_bytes, hipMemcpyHostToDevice, st));
        SPJ_TRY(hipEventRecord(desc_ev_, st));
        desc_pending_ = true;
    }
    char *dd = desc_.as<char>();
    const TileRun *d_runs = reinterpret_cast<const TileRun *>(dd + o_runs);
    const TileDesc *d_tiles = tiles_.as<TileDesc>();
    const ScanSeg *d_segs = reinterpret_cast<const ScanSeg *>(dd + o_segs);
    const HeavyJob *d_jobs = reinterpret_cast<const HeavyJob *>(dd + o_jobs);
    const CellGroup *d_cells = reinterpret_cast<const CellGroup *>(dd + o_cells);
    for (size_t i = 0; i < n; i++)   // resident tables: their partition starts
        if (specs[i].Rp_pre)
            SPJ_TRY(hipMemcpyAsync(pstart_.as<uint32_t>() + h_insts_[i].pstartR, specs[i].pst_pre,
                                   ((size_t(1) << specs[i].cfg.bits) + 1) * 4, hipMemcpyDeviceToDevice, st));

    const InstDesc *d_insts = reinterpret_cast<const InstDesc *>(dd + o_insts);
    uint32_t *ctrl = ctrl_.as<uint32_t>();
    uint32_t *out_lo = ctrl + ctrl_at(n, kCtrlLo, 0), *out_hi = ctrl + ctrl_at(n, kCtrlHi, 0);
    uint32_t *out_fill = ctrl + ctrl_at(n, kCtrlFill, 0), *err = ctrl + ctrl_at(n, kCtrlErr, 0);
    uint32_t *nheavy = ctrl + ctrl_at(n, kCtrlHeavy, 0), *spart = ctrl + ctrl_at(n, kCtrlSPart, 0);
    uint32_t *rpart = ctrl + ctrl_at(n, kCtrlRPart, 0);
    if (timed) SPJ_TRY(hipEventRecord(ev_[0], st));
    // ---- one preparation launch: control words, heavy tables and candidate
    // maps, counters, tiles, cell cursors
    const uint32_t ntiles = ntiles_;
    {
        PrepArgs pa{};
        pa.ctrl = ctrl;
        pa.nctrl = ctrl_words(n);
        pa.heavy = heavy_.as<uint32_t>();
        pa.nheavy = (uint64_t)nheavy_inst * kHeavyWords;
        pa.cand = cand_.as<uint32_t>();
        pa.ncand = (uint64_t)nheavy_inst * cand_words();
        pa.counts = counts_.as<uint32_t>();
        pa.nzero = cnt_counted;
        pa.runs = d_runs;
        pa.nruns = (uint32_t)h_runs_.size();
        pa.ntiles = ntiles;
        pa.tiles = tiles_.as<TileDesc>();
        pa.cells = d_cells;
        pa.ncells = (uint32_t)h_cells_.size();
        pa.nphj = nphj_words ? nphj_.as<unsigned long long>() : nullptr;   // cleared here: one launch fewer
        pa.nnphj = nphj_words;
        SPJ_TRY(launch_prep(pa, st));
    }

    // ---- heavy keys: two sampling launches (pre-sample, then the rest when
    // the pre-sample saw a candidate) and one selection launch for all
    if (nheavy_inst)
        SPJ_TRY(launch_heavy_detect(d_insts, d_jobs, (uint32_t)h_jobs_.size(), hblocks, cand_.as<uint32_t>(), nheavy,
                                    (uint32_t)std::max(1, device_info().cus), st));
    if (timed) SPJ_TRY(hipEventRecord(ev_[1], st));

    // ---- partition: histogram (counted tiles), scan, scatter
    SPJ_TRY(launch_hist(d_insts, d_tiles + hist_first, n_counted, counts_.as<uint32_t>(), max_bits_, st));
    // phases that launch nothing get no event (an event between two others
    // still costs the device a few microseconds)
    hist_timed_ = timed && n_counted != 0;
    scan_timed_ = timed && !h_segs_.empty();
    if (hist_timed_) SPJ_TRY(hipEventRecord(ev_[2], st));
    SPJ_TRY(launch_scan(d_segs, (uint32_t)h_segs_.size(), counts_.as<uint32_t>(), pstart_.as<uint32_t>(), max_bits_, st));
    if (scan_timed_) SPJ_TRY(hipEventRecord(ev_[3], st));
    // one persistent 1024-thread workgroup per CU (LDS-bound); R first
    // when its tiles record heavy rids for the S tiles.  (A launch of few
    // tiles keeps the full grid too: each workgroup takes three tiles of its
    // group's queue ahead, so a small grid leaves a few workgroups with all
    // the tiles.)
    const uint32_t cus = (uint32_t)std::max(1, device_info().cus);
    if (scat_split) {
        SPJ_TRY(launch_scatter(d_insts, d_tiles, scat_split, counts_.as<uint32_t>(), out_hi, err, ctrl + 4, max_bits_,
                               max_ipt_, cus, st));
        if (timed) SPJ_TRY(hipEventRecord(ev_[7], st));
    }
    split_timed_ = timed && scat_split != 0;
    SPJ_TRY(launch_scatter(d_insts, d_tiles + scat_split, ntiles - scat_split, counts_.as<uint32_t>(), out_hi, err,
                           ctrl + 4 + kScatterGroups, max_bits_, max_ipt_, cus, st));
    if (timed) SPJ_TRY(hipEventRecord(ev_[4], st));

    // ---- join (none after a first pass of two)
    partition_only_ = n && specs[0].partition_only;
    if (partition_only_) {
        if (timed) {
            SPJ_TRY(hipEventRecord(ev_[5], st));
            SPJ_TRY(hipEventRecord(ev_[6], st));
        }
        specs_n_ = n;
        h_ns_.resize(n);
        for (size_t i = 0; i < n; i++) h_ns_[i] = specs[i].nS;
        return hipSuccess;
    }
    // work items and the partitioned join, unless every instance takes the
    // no-partition join (then neither has anything to do)
    if (n_joined)
        SPJ_TRY(launch_make_items(d_insts, (uint32_t)n, pstart_.as<uint32_t>(), d_cells, counts_.as<uint32_t>(),
                                  smax, items_.as<WorkItem>(), ctrl, spart, rpart, st));
    if (timed) SPJ_TRY(hipEventRecord(ev_[5], st));
    // one persistent 1024-thread workgroup per CU (LDS-bound), also for few
    // items: a workgroup takes two items at its start, and a grid of as
    // many workgroups as items left half of them idle (2^16 x 2^18: 4
    // items, k_join 59.5 us with 256 workgroups, 109.8 us with 7)
    const uint32_t grid = cus;
    if (record_ext_) SPJ_TRY(ext_.ensure(items_bound * sizeof(uint4)));
    bool range = false;   // a sort-merge instance: the join kernel with key-range buckets
    for (const InstDesc &d : h_insts_) range |= d.alg == 2;
    if (n_joined)
        SPJ_TRY(launch_join(d_insts, items_.as<WorkItem>(), ctrl, out_lo, out_fill, err, holes_.as<uint4>(),
                            max_holes_, record_ext_ ? ext_.as<uint4>() : nullptr, grid, range, st));
    SPJ_TRY(launch_nphj(d_insts, reinterpret_cast<const NphjJob *>(dd + o_nphj), (uint32_t)h_nphj_.size(),
                        nphj_.as<uint64_t>(), nphj_words, (uint32_t)nphj_bb, (uint32_t)nphj_pb, out_lo, out_fill,
                        err, false, st));
    if (timed) SPJ_TRY(hipEventRecord(ev_[6], st));
    specs_n_ = n;
    h_ns_.resize(n);
    for (size_t i = 0; i < n; i++) h_ns_[i] = specs[i].nS;
    return hipSuccess;
}

// Make instance i's output dense when it is not already: remove the holes
// that items with unmatched S tuples left in [0, lo) and move the heavy pairs
// from [nS - hi, nS) down behind the item pairs.  Only runs when some S tuple
// found no R partner (never in a PK-FK join whose S keys all occur in R).
hipError_t Context::compact(size_t i, uint32_t lo, uint32_t hi, uint32_t fill, const std::vector<uint4> &holes)
{
    const InstDesc &d = h_insts_[i];
    std::vector<uint4> hs;
    for (const uint4 &h : holes)
        if (h.x == i) hs.push_back(h);
    std::sort(hs.begin(), hs.end(), [](const uint4 &a, const uint4 &b) { return a.y < b.y; });
    std::vector<uint4> segs;   // (src, dst, len)
    uint32_t pos = 0, dst = 0;
    for (const uint4 &h : hs) {
        if (h.y > pos) segs.push_back(make_uint4(pos, dst, h.y - pos, 0));
        dst += h.y > pos ? h.y - pos : 0;
        pos = h.y + h.z;
    }
    if (lo > pos) {
        segs.push_back(make_uint4(pos, dst, lo - pos, 0));
        dst += lo - pos;
    }
    if (dst != fill) return hipErrorUnknown;   // hole records incomplete
    if (hi) segs.push_back(make_uint4(d.nS - hi, fill, hi, 0));
    const uint64_t total = (uint64_t)fill + hi;
    SPJ_TRY(tmp_.ensure(total * 8 + 8));
    SPJ_TRY(segbuf_.ensure(segs.size() * sizeof(uint4) + 16));
    SPJ_TRY(hipMemcpyAsync(segbuf_.p, segs.data(), segs.size() * sizeof(uint4), hipMemcpyHostToDevice, stream_));
    SPJ_TRY(launch_gather_segments(d.out, tmp_.as<uint2>(), segbuf_.as<uint4>(), (uint32_t)segs.size(), stream_));
    SPJ_TRY(hipMemcpyAsync(d.out, tmp_.p, total * 8, hipMemcpyDeviceToDevice, stream_));
    return hipStreamSynchronize(stream_);
}

hipError_t Context::partition_segments(size_t inst, const uint2 **rp, const uint2 **sp, uint32_t &ng_r,
                                       std::vector<uint2> &segs_r, uint32_t &ng_s, std::vector<uint2> &segs_s)
{
    if (inst >= h_insts_.size() || !partition_only_) return hipErrorInvalidValue;
    const InstDesc &d = h_insts_[inst];
    const uint32_t F = 1u << d.bits;
    for (uint32_t rel = 0; rel < 2; rel++) {
        const uint32_t cl = rel ? d.s_cells : d.r_cells;
        uint32_t &ng = rel ? ng_s : ng_r;
        std::vector<uint2> &segs = rel ? segs_s : segs_r;
        if (!cl) {   // counted: partition p is [pstart[p], pstart[p + 1])
            std::vector<uint32_t> ps(F + 1);
            SPJ_TRY(hipMemcpyAsync(ps.data(), pstart_.as<uint32_t>() + (rel ? d.pstartS : d.pstartR), (F + 1) * 4,
                                   hipMemcpyDeviceToHost, stream_));
            SPJ_TRY(hipStreamSynchronize(stream_));
            ng = 1;
            segs.resize(F);
            for (uint32_t p = 0; p < F; p++) segs[p] = make_uint2(ps[p], ps[p + 1] - ps[p]);
            continue;
        }
        // cells: group g's cell of digit p starts at cell0 + p cap; its cursor
        // (counts[cnt_off + p], one array of F per group, the groups
        // adjacent) ends where the group's tuples of p end
        const CellGroup &g0 = h_cells_[cl - 1];
        std::vector<uint32_t> cur((size_t)kScatterGroups * F);
        SPJ_TRY(hipMemcpyAsync(cur.data(), counts_.as<uint32_t>() + g0.cnt_off, cur.size() * 4, hipMemcpyDeviceToHost,
                               stream_));
        SPJ_TRY(hipStreamSynchronize(stream_));
        ng = kScatterGroups;
        segs.resize((size_t)F * kScatterGroups);
        for (uint32_t g = 0; g < kScatterGroups; g++) {
            const CellGroup &cg = h_cells_[cl - 1 + g];
            if (cg.cnt_off != g0.cnt_off + (uint64_t)g * F) return hipErrorInvalidValue;
            for (uint32_t p = 0; p < F; p++) {
                const uint32_t st = cg.cell0 + p * cg.cap, e = cur[(size_t)g * F + p];
                if (e < st || e - st > cg.cap) return hipErrorInvalidValue;   // (an overflow redoes the batch counted)
                segs[(size_t)p * kScatterGroups + g] = make_uint2(st, e - st);
            }
        }
    }
    *rp = d.Rp;
    *sp = d.Sp;
    return hipSuccess;
}

hipError_t Context::gather(uint2 *out, const std::vector<uint4> &segs, uint64_t total)
{
    if (segs.empty()) return hipSuccess;
    SPJ_TRY(tmp_.ensure(total * 8 + 8));
    SPJ_TRY(segbuf_.ensure(segs.size() * sizeof(uint4) + 16));
    SPJ_TRY(hipMemcpyAsync(segbuf_.p, segs.data(), segs.size() * sizeof(uint4), hipMemcpyHostToDevice, stream_));
    SPJ_TRY(launch_gather_segments(out, tmp_.as<uint2>(), segbuf_.as<uint4>(), (uint32_t)segs.size(), stream_));
    SPJ_TRY(hipMemcpyAsync(out, tmp_.p, total * 8, hipMemcpyDeviceToDevice, stream_));
    return hipStreamSynchronize(stream_);
}

hipError_t Context::collect(size_t n, std::vector<uint32_t> &counts, std::vector<uint32_t> &errs, PhaseTimes *timed)
{
    h_ctrl_.resize(ctrl_words(n));
    SPJ_TRY(pin_ctrl_.ensure(ctrl_words(n) * 4));
    SPJ_TRY(hipMemcpyAsync(pin_ctrl_.p, ctrl_.p, ctrl_words(n) * 4, hipMemcpyDeviceToHost, stream_));
    SPJ_TRY(hipStreamSynchronize(stream_));
    desc_pending_ = false;
    memcpy(h_ctrl_.data(), pin_ctrl_.p, ctrl_words(n) * 4);
    counts.resize(n);
    errs.assign(h_ctrl_.begin() + ctrl_at(n, kCtrlErr, 0), h_ctrl_.begin() + ctrl_at(n, kCtrlErr, 0) + n);
    // an S cell overflowed (keys too frequent for their cells that the heavy
    // detection did not take): redo the batch with counted partitions
    if (!force_counted_ && n == h_specs_.size() &&
        std::any_of(errs.begin(), errs.end(), [](uint32_t e) { return (e & kErrCellOverflow) != 0; })) {
        force_counted_ = true;
        hipError_t e = enqueue(h_specs_, timed);
        if (e == hipSuccess) e = collect(n, counts, errs, timed);
        force_counted_ = false;
        return e;
    }
    std::vector<uint4> holes;
    bool holes_read = false;
    for (size_t i = 0; i < n; i++) {
        const uint32_t lo = h_ctrl_[ctrl_at(n, kCtrlLo, i)], hi = h_ctrl_[ctrl_at(n, kCtrlHi, i)];
        const uint32_t fill = h_ctrl_[ctrl_at(n, kCtrlFill, i)];
        counts[i] = fill + hi;
        const bool dense = (fill == lo) && (hi == 0 || (uint64_t)lo + hi == h_insts_[i].nS);
        // a first pass of two leaves its heavy pairs at the back: the
        // second pass writes the front
        if (dense || errs[i] || partition_only_) continue;
        if (fill != lo && !holes_read) {
            const uint32_t nh = h_ctrl_[2];
            if (nh > max_holes_) {
                errs[i] |= kErrOutOverflow;
                continue;
            }
            holes.resize(nh);
            if (nh) SPJ_TRY(hipMemcpy(holes.data(), holes_.p, nh * sizeof(uint4), hipMemcpyDeviceToHost));
            holes_read = true;
        }
        SPJ_TRY(compact(i, lo, hi, fill, holes));
    }
    if (timed) {
        // boundaries of the phases; a phase that launched nothing ends where it began
        hipEvent_t b[7] = {ev_[0], ev_[1], ev_[2], ev_[3], ev_[4], ev_[5], ev_[6]};
        if (!hist_timed_) b[2] = b[1];
        if (!scan_timed_) b[3] = b[2];
        float ms[6];
        for (int i = 0; i < 6; i++) {
            ms[i] = 0.f;
            if (b[i] != b[i + 1]) SPJ_TRY(hipEventElapsedTime(&ms[i], b[i], b[i + 1]));
        }
        timed->heavy = ms[0];
        timed->hist = ms[1];
        timed->scan = ms[2];
        timed->scatter = ms[3];
        timed->items = ms[4];
        timed->join = ms[5];
        SPJ_TRY(hipEventElapsedTime(&timed->total, ev_[0], ev_[6]));
        timed->scatter_r = -1;
        if (split_timed_) SPJ_TRY(hipEventElapsedTime(&timed->scatter_r, b[3], ev_[7]));
        if (n) {
            timed->heavy_keys = h_ctrl_[ctrl_at(n, kCtrlHeavy, 0)];
            const InstDesc &d0 = h_insts_[0];
            timed->r_parted = h_ctrl_[ctrl_at(n, kCtrlRPart, 0)];
            timed->s_parted = h_ctrl_[ctrl_at(n, kCtrlSPart, 0)];
            timed->s_cells = d0.s_cells != 0 ? 1u : 0u;
            timed->r_cells = d0.r_cells != 0 ? 1u : 0u;
        }
    }
    return hipSuccess;
}

void Context::set_record_extents(bool on)
{
    if (on == record_ext_) return;
    record_ext_ = on;
    drop_graph();   // the launch sequence changes
    last_valid_ = false;
}

hipError_t Context::item_extents(std::vector<uint4> &ext)
{
    ext.clear();
    if (!record_ext_ || h_ctrl_.empty()) return hipErrorInvalidValue;
    const uint32_t n = h_ctrl_[0];   // work items of the last collected batch
    ext.resize(n);
    if (n) SPJ_TRY(hipMemcpy(ext.data(), ext_.p, (size_t)n * sizeof(uint4), hipMemcpyDeviceToHost));
    return hipSuccess;
}

hipError_t Context::export_partitioned_r(size_t i, uint2 *rp_dst, uint32_t *pstart_dst)
{
    if (i >= h_insts_.size() || h_insts_[i].r_cells) return hipErrorInvalidValue;
    const InstDesc &d = h_insts_[i];
    const size_t F = size_t(1) << d.bits;
    if (d.nR) SPJ_TRY(hipMemcpyAsync(rp_dst, d.Rp, (size_t)d.nR * 8, hipMemcpyDeviceToDevice, stream_));
    SPJ_TRY(hipMemcpyAsync(pstart_dst, pstart_.as<uint32_t>() + d.pstartR, (F + 1) * 4, hipMemcpyDeviceToDevice,
                           stream_));
    return hipStreamSynchronize(stream_);
}

hipError_t Context::run(const std::vector<InstSpec> &specs, std::vector<uint32_t> &counts, std::vector<uint32_t> &errs,
                        PhaseTimes *timed)
{
    SPJ_TRY(enqueue(specs, timed));
    return collect(specs.size(), counts, errs, timed);
}

}  // namespace spj
