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

#include <memory>
#include <string>
#include <vector>

#include "spidjoin.h"
#include "spj_capi_internal.hpp"
#include "spj_engine.hpp"

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
};

using namespace spj;

static Resolved design_params(const spj_instance *in, const spj_join_design_t *d)
{
    spj_config_t c;
    spj_config_default(&c);
    if (d) {
        if (d->join_alg == 2) c.algorithm = SPJ_ALG_SMJ;
        else if (d->join_alg == 3) c.algorithm = SPJ_ALG_NL;
        else c.algorithm = SPJ_ALG_PHJ;   // 0 phj; 1 nphj is served by the partitioned hash join
        // SPID's replication switches (bank_set / rank_set > 1, pipeline.cc:251-279)
        // select the skew-resistant path: heavy-key handling.
        if (d->bank_set > 1 || d->rank_set > 1) c.heavy = 1;
    }
    return resolve(&c, (uint64_t)in->nR, (uint64_t)in->nS, -1.0);
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
    int rc = check_resolved(in->spec.cfg);
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
    hipStream_t st = in->ctx.stream();
    hipError_t e = hipSuccess;
    if (in->nR) e = hipMemcpyAsync(in->dR.p, in->hR, (size_t)in->nR * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && in->nS) e = hipMemcpyAsync(in->dS.p, in->hS, (size_t)in->nS * 8, hipMemcpyHostToDevice, st);
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

int spj_finish(spj_instance_t *in, spj_pair_t **pairs, uint64_t *count)
{
    if (!in || !pairs || !count) return fail(SPJ_EINVAL, "null argument");
    if (!in->started) return fail(SPJ_ESTATE, "FinishJoin before StartJoin");
    in->started = false;
    std::vector<uint32_t> counts, errs;
    hipError_t e = in->ctx.collect(1, counts, errs, &in->times);
    if (e != hipSuccess) return hip_fail(e, "FinishJoin");
    in->have_times = true;
    int rc = err_from_flags(errs[0], 0);
    if (rc) return rc;
    uint64_t n = counts[0];
    // aligned_alloc like the reference's result buffers (queryHostFunc.cc:1089)
    size_t bytes = ((n * 8 + 63) / 64) * 64;
    spj_pair_t *buf = (spj_pair_t *)aligned_alloc(64, bytes ? bytes : 64);
    if (!buf) return fail(SPJ_ENOMEM, "result buffer");
    if (n) {
        e = hipMemcpy(buf, in->dOut.p, n * 8, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            free(buf);
            return hip_fail(e, "FinishJoin D2H");
        }
    }
    *pairs = buf;
    *count = n;
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

void spj_free_result(spj_pair_t *pairs) { free(pairs); }

void spj_destroy(spj_instance_t *in)
{
    if (!in) return;
    if (in->started) (void)hipStreamSynchronize(in->ctx.stream());
    in->dR.release();
    in->dS.release();
    in->dOut.release();
    delete in;
}

}  // extern "C"
