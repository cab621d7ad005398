// spj_stage.cpp -- the page-locked staging ring and its copy threads.
//
// One ring per process (kSlots chunks of kChunk bytes, allocated on first
// use with hipHostMalloc), shared under a mutex: concurrent callers share the
// one host link anyway.  Each chunk's host-side copy is split over the
// worker threads; chunk c + 1 is copied on the host while the DMA engine
// moves chunk c.  A slot is reused once the event recorded after its last
// DMA has completed.
#include "spj_stage.hpp"

#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>

#include <algorithm>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace spj {

namespace {

constexpr size_t kChunk = size_t(32) << 20;
constexpr int kSlots = 4;

// Fixed team of threads running one split copy at a time: part i of
// [0, n) goes to thread i (the caller's thread takes part 0).
class CopyTeam {
  public:
    CopyTeam()
    {
        const unsigned hw = std::thread::hardware_concurrency();
        nt_ = (int)std::max(1u, std::min(8u, hw ? hw / 2 : 1u));
        for (int i = 1; i < nt_; i++) th_.emplace_back([this, i] { loop(i); });
    }
    ~CopyTeam()
    {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
            gen_++;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    void copy(void *dst, const void *src, size_t n)
    {
        if (nt_ == 1 || n < (size_t(1) << 20)) {
            memcpy(dst, src, n);
            return;
        }
        {
            std::lock_guard<std::mutex> g(m_);
            dst_ = (char *)dst;
            src_ = (const char *)src;
            n_ = n;
            left_ = nt_ - 1;
            gen_++;
        }
        cv_.notify_all();
        part(0);
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [this] { return left_ == 0; });
    }

  private:
    int nt_ = 1;
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    bool stop_ = false;
    unsigned long gen_ = 0;
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t n_ = 0;
    int left_ = 0;

    void part(int i)
    {
        // parts in 4 KiB multiples
        const size_t per = ((n_ / nt_) + 4095) & ~size_t(4095);
        const size_t b = std::min(n_, per * (size_t)i), e = std::min(n_, b + per);
        if (e > b) memcpy(dst_ + b, src_ + b, e - b);
    }
    void loop(int i)
    {
        unsigned long seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                if (stop_) return;
            }
            part(i);
            {
                std::lock_guard<std::mutex> g(m_);
                if (--left_ == 0) done_.notify_one();
            }
        }
    }
};

struct Ring {
    std::mutex m;
    char *slot[kSlots] = {};
    hipEvent_t ev[kSlots] = {};
    bool used[kSlots] = {};
    CopyTeam team;

    hipError_t ready()
    {
        for (int i = 0; i < kSlots; i++) {
            if (!slot[i]) {
                hipError_t e = hipHostMalloc((void **)&slot[i], kChunk, hipHostMallocDefault);
                if (e != hipSuccess) {
                    slot[i] = nullptr;
                    return e;
                }
            }
            if (!ev[i]) {
                hipError_t e = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming);
                if (e != hipSuccess) {
                    ev[i] = nullptr;
                    return e;
                }
            }
        }
        return hipSuccess;
    }
    // the slot's previous DMA is done
    hipError_t drain(int s)
    {
        if (!used[s]) return hipSuccess;
        used[s] = false;
        return hipEventSynchronize(ev[s]);
    }
};

Ring &ring()
{
    static Ring *r = new Ring();   // lives until the process ends (its threads with it)
    return *r;
}

}  // namespace

hipError_t stage_h2d(void *dst_dev, const void *src_host, size_t n, hipStream_t st)
{
    if (!n) return hipSuccess;
    Ring &R = ring();
    std::lock_guard<std::mutex> g(R.m);
    hipError_t e = R.ready();
    if (e != hipSuccess) return e;
    size_t off = 0;
    for (int c = 0; off < n; c++) {
        const int s = c % kSlots;
        const size_t len = std::min(kChunk, n - off);
        e = R.drain(s);
        if (e != hipSuccess) return e;
        R.team.copy(R.slot[s], (const char *)src_host + off, len);
        e = hipMemcpyAsync((char *)dst_dev + off, R.slot[s], len, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipEventRecord(R.ev[s], st);
        if (e != hipSuccess) return e;
        R.used[s] = true;
        off += len;
    }
    return hipSuccess;
}

hipError_t stage_d2h(void *dst_host, const void *src_dev, size_t n, hipStream_t st)
{
    if (!n) return hipSuccess;
    Ring &R = ring();
    std::lock_guard<std::mutex> g(R.m);
    hipError_t e = R.ready();
    if (e != hipSuccess) return e;
    for (int s = 0; s < kSlots; s++) {
        e = R.drain(s);   // slots an earlier H2D still reads
        if (e != hipSuccess) return e;
    }
    const size_t nch = (n + kChunk - 1) / kChunk;
    auto issue = [&](size_t c) -> hipError_t {
        const int s = (int)(c % kSlots);
        const size_t off = c * kChunk, len = std::min(kChunk, n - off);
        hipError_t x = hipMemcpyAsync(R.slot[s], (const char *)src_dev + off, len, hipMemcpyDeviceToHost, st);
        if (x == hipSuccess) x = hipEventRecord(R.ev[s], st);
        if (x == hipSuccess) R.used[s] = true;
        return x;
    };
    for (size_t c = 0; c < nch && c < (size_t)kSlots; c++) {
        e = issue(c);
        if (e != hipSuccess) return e;
    }
    for (size_t c = 0; c < nch; c++) {
        const int s = (int)(c % kSlots);
        const size_t off = c * kChunk, len = std::min(kChunk, n - off);
        e = R.drain(s);
        if (e != hipSuccess) return e;
        R.team.copy((char *)dst_host + off, R.slot[s], len);
        if (c + kSlots < nch) {
            e = issue(c + kSlots);
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

void *alloc_result_buffer(size_t bytes)
{
    constexpr size_t kHuge = size_t(2) << 20;
    if (bytes < kHuge) return aligned_alloc(64, ((bytes + 63) / 64) * 64 + (bytes ? 0 : 64));
    const size_t n = (bytes + kHuge - 1) / kHuge * kHuge;
    void *p = aligned_alloc(kHuge, n);
    if (p) (void)madvise(p, n, MADV_HUGEPAGE);   // advice only: without THP the pages stay 4 KiB
    return p;
}

}  // namespace spj
