// spj_serve.cpp -- serving loop: resident dimension tables, probe requests
// admitted by an HBM budget and continuously batched onto execution slots.
//
// Each table is radix-partitioned once (the R side of the join) and stays in
// HBM.  A request is an S batch against one table.  The scheduler thread
// admits queued requests in arrival order while their HBM footprint fits the
// budget, packs every admitted request into one batch for a free slot, and
// hands it to that slot's worker.  A slot is an engine Context (its own HIP
// stream and work buffers) plus a worker thread: it uploads the S batches,
// runs the batched join against the resident partitions, copies each result
// back and completes the requests.  With several slots, batches formed from
// later arrivals run while earlier ones are still on the device.
//
// The reference has no serving layer; the per-request join is the same plan
// (partition S, probe the R partitions) as query.cc's JoinInstance.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "spidjoin.h"
#include "spj_capi_internal.hpp"
#include "spj_stage.hpp"
#include "spj_engine.hpp"

namespace {

using Clock = std::chrono::steady_clock;

struct Table {
    spj::DevBuf rp, pstart;
    uint32_t nR = 0;
    spj::Resolved cfg;
};

struct Request {
    uint64_t id = 0;
    int table = 0;
    const spj_tuple_t *S = nullptr;
    uint64_t nS = 0;
    size_t footprint = 0;
    Clock::time_point t_submit, t_done;
    spj_pair_t *pairs = nullptr;
    uint64_t count = 0;
    int rc = SPJ_OK;
    bool done = false;
};

struct Slot {
    std::unique_ptr<spj::Context> ctx;
    spj::DevBuf S, out;
    std::vector<Request *> batch;
    bool busy = false;
    std::thread worker;
    std::condition_variable cv;
};

}  // namespace

struct spj_server {
    std::mutex mu;
    std::condition_variable cv_sched, cv_done;
    std::vector<std::unique_ptr<Table>> tables;
    std::deque<Request *> pending;
    std::map<uint64_t, std::unique_ptr<Request>> reqs;
    std::vector<std::unique_ptr<Slot>> slots;
    std::thread sched;
    uint64_t next_id = 1;
    size_t budget = 0, inflight = 0;
    bool stop = false;
    int device = 0;
    // stats
    uint64_t batches = 0, batched_requests = 0;
};

namespace {

// HBM held by one request while in flight: S, its partitioned copy, the
// pairs, and the partition-pass counters (bounded by S's size).
size_t footprint_of(uint64_t nS) { return (size_t)nS * 8 * 4 + (1u << 20); }

void run_batch(spj_server *sv, Slot *sl)
{
    std::vector<Request *> &b = sl->batch;
    uint64_t totS = 0;
    for (Request *r : b) totS += r->nS;
    int rc = SPJ_OK;
    hipError_t e = sl->S.ensure(totS * 8 + 8);
    if (e == hipSuccess) e = sl->out.ensure(totS * 8 + 8);
    std::vector<spj::InstSpec> specs;
    uint64_t off = 0;
    hipStream_t st = sl->ctx->stream();
    for (Request *r : b) {
        if (e != hipSuccess) break;
        const Table &t = *sv->tables[(size_t)r->table];
        e = spj::stage_h2d(sl->S.as<uint2>() + off, r->S, r->nS * 8, st);
        spj::InstSpec s;
        s.R = nullptr;
        s.nR = t.nR;
        s.S = sl->S.as<uint2>() + off;
        s.nS = (uint32_t)r->nS;
        s.out = sl->out.as<uint2>() + off;
        s.cfg = t.cfg;
        s.cfg.heavy = 0;
        s.Rp_pre = t.rp.as<uint2>();
        s.pst_pre = t.pstart.as<uint32_t>();
        specs.push_back(s);
        off += r->nS;
    }
    std::vector<uint32_t> counts, errs;
    if (e == hipSuccess) e = sl->ctx->run(specs, counts, errs, nullptr);
    if (e != hipSuccess) rc = spj::hip_fail(e, "serving batch");
    off = 0;
    for (size_t i = 0; i < b.size(); i++) {
        Request *r = b[i];
        r->rc = rc;
        if (rc == SPJ_OK) r->rc = spj::err_from_flags(errs[i], i);
        if (r->rc == SPJ_OK) {
            r->count = counts[i];
            r->pairs = (spj_pair_t *)spj::alloc_result_buffer(r->count * 8);
            if (!r->pairs) {
                r->rc = SPJ_ENOMEM;
            } else if (r->count) {
                hipError_t e2 = spj::stage_d2h(r->pairs, sl->out.as<uint2>() + off, r->count * 8, st);
                if (e2 != hipSuccess) r->rc = spj::hip_fail(e2, "serving D2H");
            }
        }
        off += r->nS;
    }
    std::lock_guard<std::mutex> lk(sv->mu);
    for (Request *r : b) {
        r->t_done = Clock::now();
        r->done = true;
        sv->inflight -= r->footprint;
    }
    sv->batches++;
    sv->batched_requests += b.size();
    b.clear();
    sl->busy = false;
    sv->cv_done.notify_all();
    sv->cv_sched.notify_all();
}

void worker_loop(spj_server *sv, Slot *sl)
{
    (void)hipSetDevice(sv->device);
    std::unique_lock<std::mutex> lk(sv->mu);
    for (;;) {
        sl->cv.wait(lk, [&] { return sv->stop || !sl->batch.empty(); });
        if (sl->batch.empty() && sv->stop) return;
        lk.unlock();
        run_batch(sv, sl);
        lk.lock();
    }
}

void sched_loop(spj_server *sv)
{
    std::unique_lock<std::mutex> lk(sv->mu);
    for (;;) {
        sv->cv_sched.wait(lk, [&] {
            if (sv->stop) return true;
            if (sv->pending.empty()) return false;
            for (auto &s : sv->slots)
                if (!s->busy) return true;
            return false;
        });
        if (sv->stop) return;
        Slot *free_slot = nullptr;
        for (auto &s : sv->slots)
            if (!s->busy) {
                free_slot = s.get();
                break;
            }
        // admit in arrival order while the footprint fits the budget
        while (!sv->pending.empty()) {
            Request *r = sv->pending.front();
            if (sv->inflight + r->footprint > sv->budget && !free_slot->batch.empty()) break;
            if (sv->inflight + r->footprint > sv->budget && sv->inflight > 0) break;   // wait for memory
            sv->pending.pop_front();
            sv->inflight += r->footprint;
            free_slot->batch.push_back(r);
        }
        if (free_slot->batch.empty()) {
            // nothing fits until a batch finishes
            sv->cv_sched.wait(lk);
            continue;
        }
        free_slot->busy = true;
        free_slot->cv.notify_one();
    }
}

}  // namespace

extern "C" {

spj_server_t *spj_server_create(uint64_t hbm_budget_bytes, int slots)
{
    if (slots < 1 || slots > 16 || hbm_budget_bytes == 0) {
        spj::fail(SPJ_EINVAL, "slots must be 1..16 and the budget non-zero");
        return nullptr;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        spj::fail(SPJ_EHIP, "no HIP device");
        return nullptr;
    }
    spj_server *sv = new spj_server();
    sv->budget = hbm_budget_bytes;
    sv->device = dev;
    for (int i = 0; i < slots; i++) {
        auto s = std::make_unique<Slot>();
        s->ctx = std::make_unique<spj::Context>();
        sv->slots.push_back(std::move(s));
    }
    for (auto &s : sv->slots) s->worker = std::thread(worker_loop, sv, s.get());
    sv->sched = std::thread(sched_loop, sv);
    return sv;
}

int spj_server_add_table(spj_server_t *sv, const spj_tuple_t *R, uint64_t nR, int *table_id)
{
    if (!sv || (!R && nR) || !table_id) return spj::fail(SPJ_EINVAL, "null argument");
    int rc = spj::validate_sizes(nR, 0);
    if (rc) return rc;
    auto t = std::make_unique<Table>();
    t->nR = (uint32_t)nR;
    t->cfg = spj::resolve(nullptr, nR, nR, -1.0);
    t->cfg.heavy = 0;
    // partition R once through the engine (an instance with an empty S)
    spj::Context ctx;
    spj::DevBuf dR, dummy;
    hipError_t e = dR.ensure(nR * 8 + 8);
    if (e == hipSuccess) e = dummy.ensure(64);
    if (e == hipSuccess) e = spj::stage_h2d(dR.p, R, nR * 8, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    spj::InstSpec s;
    s.R = dR.as<uint2>();
    s.nR = (uint32_t)nR;
    s.S = dummy.as<uint2>();
    s.nS = 0;
    s.out = dummy.as<uint2>();
    s.cfg = t->cfg;
    s.r_contiguous = true;   // exported as one run per partition
    std::vector<uint32_t> counts, errs;
    if (e == hipSuccess) e = ctx.run({s}, counts, errs, nullptr);
    if (e == hipSuccess) e = t->rp.ensure(nR * 8 + 8);
    if (e == hipSuccess) e = t->pstart.ensure(((size_t(1) << t->cfg.bits) + 1) * 4);
    if (e == hipSuccess) e = ctx.export_partitioned_r(0, t->rp.as<uint2>(), t->pstart.as<uint32_t>());
    dR.release();
    dummy.release();
    if (e != hipSuccess) return spj::hip_fail(e, "add_table");
    std::lock_guard<std::mutex> lk(sv->mu);
    *table_id = (int)sv->tables.size();
    sv->tables.push_back(std::move(t));
    return SPJ_OK;
}

int spj_server_submit(spj_server_t *sv, int table_id, const spj_tuple_t *S, uint64_t nS, uint64_t *request_id)
{
    if (!sv || (!S && nS) || !request_id) return spj::fail(SPJ_EINVAL, "null argument");
    int rc = spj::validate_sizes(0, nS);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(sv->mu);
    if (table_id < 0 || (size_t)table_id >= sv->tables.size()) return spj::fail(SPJ_EINVAL, "no table %d", table_id);
    const size_t fp = footprint_of(nS);
    if (fp > sv->budget) return spj::fail(SPJ_EBUSY, "request needs %zu bytes, budget %zu", fp, sv->budget);
    auto r = std::make_unique<Request>();
    r->id = sv->next_id++;
    r->table = table_id;
    r->S = S;
    r->nS = nS;
    r->footprint = fp;
    r->t_submit = Clock::now();
    *request_id = r->id;
    sv->pending.push_back(r.get());
    sv->reqs[r->id] = std::move(r);
    sv->cv_sched.notify_all();
    return SPJ_OK;
}

int spj_server_wait(spj_server_t *sv, uint64_t id, spj_pair_t **pairs, uint64_t *count, double *latency_ms)
{
    if (!sv || !pairs || !count) return spj::fail(SPJ_EINVAL, "null argument");
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return spj::fail(SPJ_EINVAL, "unknown request %llu", (unsigned long long)id);
    Request *r = it->second.get();
    sv->cv_done.wait(lk, [&] { return r->done; });
    const int rc = r->rc;
    *pairs = r->pairs;
    *count = r->count;
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    sv->reqs.erase(it);
    if (rc) return spj::fail(rc, "request %llu failed", (unsigned long long)id);
    return SPJ_OK;
}

int spj_server_stats(spj_server_t *sv, uint64_t *batches, uint64_t *requests)
{
    if (!sv) return spj::fail(SPJ_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(sv->mu);
    if (batches) *batches = sv->batches;
    if (requests) *requests = sv->batched_requests;
    return SPJ_OK;
}

void spj_server_destroy(spj_server_t *sv)
{
    if (!sv) return;
    {
        std::unique_lock<std::mutex> lk(sv->mu);
        sv->cv_done.wait(lk, [&] {
            if (!sv->pending.empty()) return false;
            for (auto &s : sv->slots)
                if (s->busy) return false;
            return true;
        });
        sv->stop = true;
        sv->cv_sched.notify_all();
        for (auto &s : sv->slots) s->cv.notify_all();
    }
    sv->sched.join();
    for (auto &s : sv->slots) s->worker.join();
    for (auto &kv : sv->reqs) free(kv.second->pairs);
    delete sv;
}

}  // extern "C"
