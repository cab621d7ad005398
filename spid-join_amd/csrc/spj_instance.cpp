// spj_instance.cpp -- the reference-shaped JoinInstance life cycle
// (src/query.cc: ctor 842-961, LoadColumn 213-317, PrepareJoin 178-186,
// StartJoin 188-191, FinishJoin 193-197, ExecuteJoin 199-211) on one GPU.
//
// LoadColumn records the caller's host pointer (borrowed, as the reference's
// zero-copy split at query.cc:307).  PrepareJoin allocates device buffers.
// StartJoin enqueues the H2D copies and all kernels on the instance's own
// stream and returns (asynchronous, like the reference's signal to its rank
// threads).  FinishJoin waits and copies back exactly the result pairs into
// a library-allocated host buffer that the caller owns.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "spidjoin.h"
#include "spj_capi_internal.hpp"
#include "spj_engine.hpp"
#include "spj_plan.hpp"
#include "spj_stage.hpp"

struct spj_instance {
    int ranks = 16, rank_groups = 1, ranks_per_group = 16;
    const spj_tuple_t *hR = nullptr, *hS = nullptr;
    int64_t nR = -1, nS = -1;
    bool prepared = false, started = false;
    spj::Context ctx;
    spj::DevBuf dR, dS, dOut;
    spj::InstSpec spec;
    spj::PhaseTimes times;
    bool have_times = false;
    // join plan: set explicitly, else chosen from the design at StartJoin
    spj::Plan plan, run_plan;
    bool plan_set = false;
    // timeline of the last join: CLOCK_MONOTONIC ns of the first H2D copy's
    // enqueue (t0) and around the result copy; events around the two copies
    hipEvent_t ev[3] = {};
    int64_t t0_ns = 0, recv0_ns = 0, recv1_ns = 0;
    std::vector<spj::NodeTime> node_times;
    int layout = SPJ_RESULT_EXACT;   // FinishJoin's result buffers
};

using namespace spj;

// log2 of the reference's DPU count, 64 per rank (NUM_DPU_RANK), ranks a
// power of two as the reference's rank allocation (idpHandler.hpp:134)
static uint32_t rank_dpu_bits(int ranks)
{
    uint32_t b = 6;
    while ((1 << (b - 6)) < ranks) b++;
    return b;
}

static int64_t now_ns()
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (int64_t)ts.tv_sec * 1000000000 + ts.tv_nsec;
}

// The plan StartJoin runs: the one set, else SetQueryTree's choice
// (pipeline.cc:251-279): replication (bank_set or rank_set > 1) selects the
// plan with BG_BROADCAST.
static bool design_replicates(const spj_join_design_t *d) { return d && (d->bank_set > 1 || d->rank_set > 1); }

static Resolved design_params(const spj_instance *in, const spj_join_design_t *d)
{
    spj_config_t c;
    spj_config_default(&c);
    if (d) {
        if (d->join_alg == 2) c.algorithm = SPJ_ALG_SMJ;
        else if (d->join_alg == 3) c.algorithm = SPJ_ALG_NL;
        else if (d->join_alg == 1) c.algorithm = SPJ_ALG_NPHJ;
        else c.algorithm = SPJ_ALG_PHJ;
    }
    // SPID's replication switches (bank_set / rank_set > 1, pipeline.cc:251-279)
    // select the skew-resistant plan; its BG_BROADCAST is heavy-key handling
    if (in->plan_set ? in->plan.replicate : design_replicates(d)) c.heavy = 1;
    return resolve(&c, (uint64_t)in->nR, (uint64_t)in->nS, -1.0);
}

// Start and stop of every plan node (ms from t0), from the copy events, the
// GPU phase times (back to back on the instance's stream) and the host
// clock around the result copy.
static void plan_times(spj_instance *in, float send_r, float send_s)
{
    using spj::PlanOp;
    const spj::PhaseTimes &t = in->times;
    const double c = (double)send_r + send_s;        // end of both copies
    const double h1 = c + t.heavy;                   // end of heavy-key detection
    const double p0 = in->run_plan.replicate ? h1 : c;   // without BG_BROADCAST the detection counts as partitioning
    const double pe = h1 + t.hist + t.scan + t.scatter;  // end of partitioning
    const bool split = t.scatter_r >= 0;
    const double r_end = split ? h1 + t.hist + t.scan + t.scatter_r : pe;
    const double s0 = split ? r_end : p0;
    const double je = pe + t.items + t.join;
    const double rc0 = in->recv0_ns ? (double)(in->recv0_ns - in->t0_ns) / 1e6 : je;
    const double rc1 = in->recv1_ns ? (double)(in->recv1_ns - in->t0_ns) / 1e6 : rc0;
    in->node_times.assign(in->run_plan.nodes.size(), spj::NodeTime());
    double cur = 0;
    for (size_t i = 0; i < in->run_plan.nodes.size(); i++) {
        const spj::PlanNode &n = in->run_plan.nodes[i];
        double a = cur, b = cur;
        switch (n.op) {
        case PlanOp::Send: a = n.side ? send_r : 0.0; b = n.side ? c : send_r; break;
        case PlanOp::Sync: break;
        case PlanOp::BgBroadcast: a = c; b = h1; break;
        case PlanOp::ChipwiseBroadcast: a = b = h1; break;
        case PlanOp::Partition:
        case PlanOp::ChipwisePartition: a = n.side ? s0 : p0; b = n.side ? pe : r_end; break;
        case PlanOp::RotateAndStream: a = b = pe; break;
        case PlanOp::RnsJoin:
        case PlanOp::IdpJoin: a = pe; b = je; break;
        case PlanOp::Recv: a = rc0; b = rc1; break;
        }
        in->node_times[i].start_ms = a;
        in->node_times[i].stop_ms = b;
        if (b > cur) cur = b;
    }
}

extern "C" {

void spj_join_design_default(spj_join_design_t *d)
{
    if (!d) return;
    d->input_size = 500000;
    d->rank_num = 16;
    d->zipf_factor = 0;
    d->rs_ratio = 8;
    d->rank_set = 1;
    d->bank_set = 1;
    d->join_alg = 0;
}

spj_instance_t *spj_create(int num_rank_allocated, int num_rankgroup, int num_rank_in_rankgroup)
{
    if (num_rank_allocated < 0 || num_rankgroup < 0 || num_rank_in_rankgroup < 0) {
        fail(SPJ_EINVAL, "negative rank counts");
        return nullptr;
    }
    spj_instance *in = new (std::nothrow) spj_instance();
    if (!in) {
        fail(SPJ_ENOMEM, "host allocation");
        return nullptr;
    }
    in->ranks = num_rank_allocated;
    in->rank_groups = num_rankgroup;
    in->ranks_per_group = num_rank_in_rankgroup;
    return in;
}

int spj_load_column(spj_instance_t *in, const void *data, int64_t n, const char *name)
{
    if (!in || !name) return fail(SPJ_EINVAL, "null argument");
    if (n < 0 || (n > 0 && !data)) return fail(SPJ_EINVAL, "bad column");
    if (validate_sizes((uint64_t)n, 0)) return SPJ_EINVAL;
    if (in->started) return fail(SPJ_ESTATE, "LoadColumn while a join is running");
    // query.cc:215-227: only "left" (R, build) and "right" (S, probe)
    if (strcmp(name, "left") == 0) {
        in->hR = (const spj_tuple_t *)data;
        in->nR = n;
    } else if (strcmp(name, "right") == 0) {
        in->hS = (const spj_tuple_t *)data;
        in->nS = n;
    } else {
        return fail(SPJ_EINVAL, "Column name error: %s should be left or right", name);
    }
    in->prepared = false;
    return SPJ_OK;
}

int spj_prepare(spj_instance_t *in, const spj_join_design_t *d)
{
    if (!in) return fail(SPJ_EINVAL, "null instance");
    if (in->nR < 0 || in->nS < 0) return fail(SPJ_ESTATE, "PrepareJoin before both LoadColumn calls");
    if (in->started) return fail(SPJ_ESTATE, "PrepareJoin while a join is running");
    hipError_t e = in->dR.ensure((size_t)in->nR * 8 + 8);
    if (e == hipSuccess) e = in->dS.ensure((size_t)in->nS * 8 + 8);
    if (e == hipSuccess) e = in->dOut.ensure((size_t)in->nS * 8 + 8);
    if (e != hipSuccess) return hip_fail(e, "PrepareJoin allocation");
    in->spec.cfg = design_params(in, d);
    int rc = check_resolved(in->spec.cfg, (uint64_t)in->nR);
    if (rc) return rc;
    in->prepared = true;
    return SPJ_OK;
}

int spj_start(spj_instance_t *in, const spj_join_design_t *d)
{
    if (!in) return fail(SPJ_EINVAL, "null instance");
    if (!in->prepared) {
        int rc = spj_prepare(in, d);
        if (rc) return rc;
    }
    if (in->started) return fail(SPJ_ESTATE, "StartJoin twice without FinishJoin");
    in->spec.cfg = design_params(in, d);
    if (in->layout == SPJ_RESULT_RANK_PADDED) {
        // the reference's per-DPU result blocks: a pair belongs to the DPU of
        // its S tuple's global partition, glb_partition_hash(key) & (D - 1)
        // with D = 64 ranks DPUs (glb_partition_count.c:192-195) -- the low
        // bits of this build's own partition digit, given at least log2 D
        // radix bits.  Heavy pairs bypass the partitions, so heavy handling
        // is off; the replicating plans place S by position, which has no
        // counterpart here.
        if (design_replicates(d) || (in->plan_set && in->plan.replicate))
            return fail(SPJ_EINVAL, "the rank-padded result layout covers the plain plan (bank_set = rank_set = 1)");
        if (in->spec.cfg.alg == SPJ_ALG_NPHJ)
            return fail(SPJ_EINVAL, "the rank-padded result layout needs a partitioned join (join_alg 0, 2 or 3)");
        if (in->ranks < 1 || (in->ranks & (in->ranks - 1)))
            return fail(SPJ_EINVAL, "the rank-padded result layout needs a power-of-two rank count, got %d", in->ranks);
        const uint32_t dbits = rank_dpu_bits(in->ranks);
        if (dbits > kMaxBits) return fail(SPJ_EINVAL, "the rank-padded result layout takes at most 64 ranks");
        in->spec.cfg.bits = std::max(in->spec.cfg.bits, dbits);
        in->spec.cfg.heavy = 0;
    }
    in->ctx.set_record_extents(in->layout == SPJ_RESULT_RANK_PADDED);
    in->run_plan = in->plan_set ? in->plan : spj::default_plan(design_replicates(d));
    hipStream_t st = in->ctx.stream();
    hipError_t e = hipSuccess;
    for (int i = 0; i < 3 && e == hipSuccess; i++)
        if (!in->ev[i]) e = hipEventCreate(&in->ev[i]);
    if (e != hipSuccess) return hip_fail(e, "StartJoin events");
    in->t0_ns = now_ns();
    in->recv0_ns = in->recv1_ns = 0;
    // through the page-locked staging ring: returns once the columns have
    // been read (the DMA of the last chunk may still run)
    e = hipEventRecord(in->ev[0], st);
    if (e == hipSuccess) e = stage_h2d(in->dR.p, in->hR, (size_t)in->nR * 8, st);
    if (e == hipSuccess) e = hipEventRecord(in->ev[1], st);
    if (e == hipSuccess) e = stage_h2d(in->dS.p, in->hS, (size_t)in->nS * 8, st);
    if (e == hipSuccess) e = hipEventRecord(in->ev[2], st);
    if (e != hipSuccess) return hip_fail(e, "StartJoin H2D");
    in->spec.R = in->dR.as<uint2>();
    in->spec.S = in->dS.as<uint2>();
    in->spec.out = in->dOut.as<uint2>();
    in->spec.nR = (uint32_t)in->nR;
    in->spec.nS = (uint32_t)in->nS;
    e = in->ctx.enqueue({in->spec}, &in->times);
    if (e != hipSuccess) return hip_fail(e, "StartJoin enqueue");
    in->started = true;
    return SPJ_OK;
}

static int finish_exact(spj_instance *in, spj_pair_t **pairs, uint64_t *count);
static int finish_rank_padded(spj_instance *in, spj_pair_t **bufs, uint64_t *bytes, int cap, int *nbufs);

int spj_set_result_layout(spj_instance_t *in, int layout)
{
    if (!in) return fail(SPJ_EINVAL, "null instance");
    if (layout != SPJ_RESULT_EXACT && layout != SPJ_RESULT_RANK_PADDED) return fail(SPJ_EINVAL, "layout %d", layout);
    if (in->started) return fail(SPJ_ESTATE, "layout change while a join is running");
    in->layout = layout;
    return SPJ_OK;
}

int spj_finish_buffers(spj_instance_t *in, spj_pair_t **bufs, uint64_t *bytes, int cap, int *nbufs)
{
    if (!in || !bufs || !bytes || !nbufs) return fail(SPJ_EINVAL, "null argument");
    if (!in->started) return fail(SPJ_ESTATE, "FinishJoin before StartJoin");
    if (in->layout == SPJ_RESULT_RANK_PADDED) return finish_rank_padded(in, bufs, bytes, cap, nbufs);
    if (cap < 1) return fail(SPJ_EINVAL, "room for one buffer needed");
    uint64_t n = 0;
    int rc = finish_exact(in, &bufs[0], &n);
    if (rc) return rc;
    bytes[0] = n * sizeof(spj_pair_t);
    *nbufs = 1;
    return SPJ_OK;
}

int spj_finish(spj_instance_t *in, spj_pair_t **pairs, uint64_t *count)
{
    if (!in || !pairs || !count) return fail(SPJ_EINVAL, "null argument");
    if (!in->started) return fail(SPJ_ESTATE, "FinishJoin before StartJoin");
    if (in->layout != SPJ_RESULT_EXACT)
        return fail(SPJ_ESTATE, "the instance returns rank-padded buffers: use spj_finish_buffers");
    return finish_exact(in, pairs, count);
}

// Wait for the join, record its timeline, check its flags.
static int finish_wait(spj_instance *in, uint64_t *count, float *send_r, float *send_s)
{
    in->started = false;
    std::vector<uint32_t> counts, errs;
    hipError_t e = in->ctx.collect(1, counts, errs, &in->times);
    if (e != hipSuccess) return hip_fail(e, "FinishJoin");
    in->have_times = true;
    e = hipEventElapsedTime(send_r, in->ev[0], in->ev[1]);
    if (e == hipSuccess) e = hipEventElapsedTime(send_s, in->ev[1], in->ev[2]);
    if (e != hipSuccess) return hip_fail(e, "FinishJoin timing");
    plan_times(in, *send_r, *send_s);
    int rc = err_from_flags(errs[0], 0);
    if (rc) return rc;
    *count = counts[0];
    return SPJ_OK;
}

static int finish_exact(spj_instance *in, spj_pair_t **pairs, uint64_t *count)
{
    uint64_t n = 0;
    float send_r = 0, send_s = 0;
    int rc = finish_wait(in, &n, &send_r, &send_s);
    if (rc) return rc;
    hipError_t e;
    // aligned_alloc like the reference's result buffers (queryHostFunc.cc:1089)
    spj_pair_t *buf = (spj_pair_t *)alloc_result_buffer(n * 8);
    if (!buf) return fail(SPJ_ENOMEM, "result buffer");
    in->recv0_ns = now_ns();
    if (n) {
        e = stage_d2h(buf, in->dOut.p, n * 8, in->ctx.stream());
        if (e != hipSuccess) {
            free(buf);
            return hip_fail(e, "FinishJoin D2H");
        }
    }
    in->recv1_ns = now_ns();
    plan_times(in, send_r, send_s);
    *pairs = buf;
    *count = n;
    return SPJ_OK;
}

// The reference's gathered result (queryHostFunc.cc:1058-1126,
// finish_join.c:48-141): one buffer per rank of 64 DPU blocks, each block
// the DPU's pairs followed by (0, 0) pairs up to the rank's largest block;
// the byte count includes the padding.  A pair's DPU is its S tuple's
// global partition modulo D = 64 ranks (the join ran with at least log2 D
// radix bits, so every work item belongs to one DPU).  Pairs inside a block
// are in no particular order, as in the reference.
static int finish_rank_padded(spj_instance *in, spj_pair_t **bufs, uint64_t *bytes, int cap, int *nbufs)
{
    const int ranks = in->ranks;
    if (cap < ranks) return fail(SPJ_EINVAL, "room for %d buffers needed", ranks);
    uint64_t n = 0;
    float send_r = 0, send_s = 0;
    int rc = finish_wait(in, &n, &send_r, &send_s);
    if (rc) return rc;
    std::vector<uint4> ext;
    hipError_t e = in->ctx.item_extents(ext);
    if (e != hipSuccess) return hip_fail(e, "FinishJoin item extents");
    in->recv0_ns = now_ns();
    std::vector<spj_pair_t> all(n);
    if (n) {
        e = stage_d2h(all.data(), in->dOut.p, n * 8, in->ctx.stream());
        if (e != hipSuccess) return hip_fail(e, "FinishJoin D2H");
    }
    // the result was compacted: the items' pairs lie back to back in the
    // order of their reservations
    std::sort(ext.begin(), ext.end(), [](const uint4 &a, const uint4 &b) { return a.z < b.z; });
    const uint32_t D = 64u * (uint32_t)ranks;
    std::vector<uint64_t> per_dpu(D, 0);
    uint64_t pos = 0;
    for (const uint4 &x : ext) {
        per_dpu[x.y & (D - 1)] += x.w;
        pos += x.w;
    }
    if (pos != n) return fail(SPJ_EHIP, "FinishJoin: item extents cover %llu of %llu pairs", (unsigned long long)pos,
                              (unsigned long long)n);
    std::vector<uint64_t> block(ranks, 0), fill(D, 0);
    for (uint32_t d = 0; d < D; d++) block[d / 64] = std::max(block[d / 64], per_dpu[d]);
    int made = 0;
    for (int r = 0; r < ranks; r++) {
        const uint64_t b = 64 * block[r] * sizeof(spj_pair_t);
        spj_pair_t *buf = (spj_pair_t *)alloc_result_buffer(b ? b : 64);
        if (!buf) {
            for (int k = 0; k < made; k++) free(bufs[k]);
            return fail(SPJ_ENOMEM, "result buffer");
        }
        memset(buf, 0, b);   // the (0, 0) padding of finish_join.c
        bufs[r] = buf;
        bytes[r] = b;
        made++;
    }
    pos = 0;
    for (const uint4 &x : ext) {
        const uint32_t d = x.y & (D - 1), r = d / 64;
        spj_pair_t *dst = bufs[r] + (uint64_t)(d % 64) * block[r] + fill[d];
        if (x.w) memcpy(dst, all.data() + pos, (size_t)x.w * sizeof(spj_pair_t));
        fill[d] += x.w;
        pos += x.w;
    }
    in->recv1_ns = now_ns();
    plan_times(in, send_r, send_s);
    *nbufs = ranks;
    return SPJ_OK;
}

int spj_execute(spj_instance_t *in, const char *join_type, const spj_join_design_t *d, spj_pair_t **pairs,
                uint64_t *count)
{
    if (!in || !join_type) return fail(SPJ_EINVAL, "null argument");
    // query.cc:199-211: "phj" and "smj" only
    spj_join_design_t dd;
    if (d) dd = *d;
    else spj_join_design_default(&dd);
    if (strcmp(join_type, "phj") == 0) dd.join_alg = 0;
    else if (strcmp(join_type, "smj") == 0) dd.join_alg = 2;
    else return fail(SPJ_EINVAL, "Not supported join algorithm %s", join_type);
    int rc = spj_start(in, &dd);
    if (rc) return rc;
    return spj_finish(in, pairs, count);
}

int spj_last_phase_times(spj_instance_t *in, spj_phase_times_t *t)
{
    if (!in || !t) return fail(SPJ_EINVAL, "null argument");
    if (!in->have_times) return fail(SPJ_ESTATE, "no finished join");
    fill_times(in->times, t);
    return SPJ_OK;
}

int spj_set_query_plan(spj_instance_t *in, const char *plan_json)
{
    if (!in) return fail(SPJ_EINVAL, "null instance");
    if (in->started) return fail(SPJ_ESTATE, "plan change while a join is running");
    if (!plan_json) {   // back to the design's choice
        in->plan_set = false;
        return SPJ_OK;
    }
    spj::Plan p;
    int rc = spj::parse_plan(plan_json, &p);
    if (rc) return rc;
    in->plan = std::move(p);
    in->plan_set = true;
    return SPJ_OK;
}

int spj_plan_check(const char *plan_json, int *num_nodes, int *replicate)
{
    spj::Plan p;
    int rc = spj::parse_plan(plan_json, &p);
    if (rc) return rc;
    if (num_nodes) *num_nodes = (int)p.nodes.size();
    if (replicate) *replicate = p.replicate ? 1 : 0;
    return SPJ_OK;
}

int spj_timeline(spj_instance_t *in, int64_t start_sec, int64_t start_nsec, int rank_group, char *buf, uint64_t cap,
                 uint64_t *len)
{
    if (!in || !len) return fail(SPJ_EINVAL, "null argument");
    if (!in->have_times || in->node_times.empty()) return fail(SPJ_ESTATE, "no finished join");
    const std::vector<uint64_t> parted = {in->times.r_parted, in->times.s_parted};
    const std::string s = spj::timeline_text(in->run_plan, in->node_times, in->t0_ns,
                                             start_sec * 1000000000 + start_nsec, rank_group, parted);
    *len = s.size();
    if (!buf) return SPJ_OK;
    if (cap <= s.size()) return fail(SPJ_EINVAL, "timeline needs %zu bytes", s.size() + 1);
    memcpy(buf, s.c_str(), s.size() + 1);
    return SPJ_OK;
}

void spj_free_result(spj_pair_t *pairs) { free(pairs); }

void spj_destroy(spj_instance_t *in)
{
    if (!in) return;
    if (in->started) (void)hipStreamSynchronize(in->ctx.stream());
    in->dR.release();
    in->dS.release();
    in->dOut.release();
    for (hipEvent_t e : in->ev)
        if (e) (void)hipEventDestroy(e);
    delete in;
}

}  // extern "C"
