// spj_datagen.cpp -- seeded synthetic relations for the join workloads.
//
// The reference has no generator (test/join/main.cc:30-48 fills R = {(i+1,
// i+3)} and leaves Zipf as a stub), so this is the build's own design:
//   R: keys are a Fisher-Yates permutation of 1..n (unique, never 0) drawn
//      with Lemire's bounded integers from a SplitMix64 stream; rid = row.
//   S: rank r in [0, nR) drawn Zipf(s) (P ~ (r+1)^-s) through a frozen
//      uint64 inverse-CDF table plus a guide table, or uniformly for s = 0;
//      key = R.key[sigma(r)] for an independent seeded permutation sigma of
//      R's rows (a seeded random rank order); rid = row.
// S is generated in fixed blocks of 65536 tuples, each from its own stream,
// so the bytes do not depend on the thread count.
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "spidjoin.h"

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // Lemire's nearly divisionless bounded integer in [0, range).
    uint64_t bounded(uint64_t range)
    {
        uint64_t x = next();
        __uint128_t m = (__uint128_t)x * range;
        uint64_t l = (uint64_t)m;
        if (l < range) {
            uint64_t t = (0 - range) % range;
            while (l < t) {
                x = next();
                m = (__uint128_t)x * range;
                l = (uint64_t)m;
            }
        }
        return (uint64_t)(m >> 64);
    }
};

uint64_t stream_seed(uint64_t seed, uint64_t tag, uint64_t idx)
{
    SplitMix64 g(seed ^ (tag * 0xD1B54A32D192ED03ull) ^ (idx * 0x8CB92BA72F3D8DD7ull));
    return g.next();
}

void fisher_yates(uint32_t *a, uint64_t n, uint64_t seed)
{
    SplitMix64 g(seed);
    for (uint64_t i = n; i > 1; i--) {
        uint64_t j = g.bounded(i);
        std::swap(a[i - 1], a[j]);
    }
}

int hw_threads(int req)
{
    if (req > 0) return req;
    unsigned c = std::thread::hardware_concurrency();
    return c ? (int)std::min(c, 64u) : 1;
}

template <class F>
void parallel_blocks(uint64_t nblocks, int threads, F fn)
{
    if (threads <= 1 || nblocks <= 1) {
        for (uint64_t b = 0; b < nblocks; b++) fn(b);
        return;
    }
    std::vector<std::thread> th;
    std::atomic<uint64_t> next{0};
    for (int t = 0; t < threads; t++)
        th.emplace_back([&] {
            for (;;) {
                uint64_t b = next.fetch_add(1);
                if (b >= nblocks) break;
                fn(b);
            }
        });
    for (auto &x : th) x.join();
}

}  // namespace

extern "C" int spj_gen_r(spj_tuple_t *R, uint64_t nR, uint64_t seed)
{
    if (!R && nR) return -1;
    if (nR >= 0xFFFFFFFFull) return -1;
    std::vector<uint32_t> keys(nR);
    for (uint64_t i = 0; i < nR; i++) keys[i] = (uint32_t)(i + 1);
    fisher_yates(keys.data(), nR, stream_seed(seed, 1, 0));
    for (uint64_t i = 0; i < nR; i++) {
        R[i].key = keys[i];
        R[i].rid = (uint32_t)i;
    }
    return 0;
}

extern "C" int spj_gen_s(spj_tuple_t *S, uint64_t nS, const spj_tuple_t *R, uint64_t nR, double zipf, uint64_t seed,
                         int threads)
{
    if ((!S && nS) || (!R && nR) || zipf < 0) return -1;
    if (nS == 0) return 0;
    if (nR == 0) return -1;
    threads = hw_threads(threads);
    // sigma: random rank order over R's rows
    std::vector<uint32_t> sigma(nR);
    for (uint64_t i = 0; i < nR; i++) sigma[i] = (uint32_t)i;
    fisher_yates(sigma.data(), nR, stream_seed(seed, 2, 0));

    std::vector<uint64_t> thr;
    std::vector<uint32_t> guide;
    uint32_t glog = 0;
    const bool uniform = zipf == 0.0;
    if (!uniform) {
        // Frozen inverse CDF over 63-bit uniforms: thr[r] = floor(CDF(r) * 2^63),
        // thr[n-1] = 2^63.  Partial sums in fixed 2^20 chunks (deterministic).
        thr.resize(nR);
        const uint64_t CH = 1u << 20, nch = (nR + CH - 1) / CH;
        std::vector<double> csum(nch);
        std::vector<double> w(nR);
        parallel_blocks(nch, threads, [&](uint64_t c) {
            double s = 0;
            for (uint64_t r = c * CH; r < std::min(nR, (c + 1) * CH); r++) {
                w[r] = pow((double)(r + 1), -zipf);
                s += w[r];
            }
            csum[c] = s;
        });
        std::vector<double> cbase(nch);
        double W = 0;
        for (uint64_t c = 0; c < nch; c++) {
            cbase[c] = W;
            W += csum[c];
        }
        const double scale = ldexp(1.0, 63) / W;
        parallel_blocks(nch, threads, [&](uint64_t c) {
            double s = cbase[c];
            for (uint64_t r = c * CH; r < std::min(nR, (c + 1) * CH); r++) {
                s += w[r];
                double v = s * scale;
                thr[r] = v >= ldexp(1.0, 63) ? (1ull << 63) : (uint64_t)v;
            }
        });
        thr[nR - 1] = 1ull << 63;
        for (uint64_t r = 1; r < nR; r++)   // enforce monotone (rounding)
            if (thr[r] < thr[r - 1]) thr[r] = thr[r - 1];
        glog = 1;
        while ((1ull << glog) < nR && glog < 22) glog++;
        const uint64_t G = 1ull << glog;
        guide.resize(G + 1);
        uint64_t r = 0;
        for (uint64_t g = 0; g <= G; g++) {
            uint64_t lo = (g >= G) ? (1ull << 63) : (g << (63 - glog));
            while (r < nR - 1 && thr[r] <= lo) r++;
            guide[g] = (uint32_t)r;
        }
    }
    const uint64_t BS = 1u << 16, nblk = (nS + BS - 1) / BS;
    parallel_blocks(nblk, threads, [&](uint64_t b) {
        SplitMix64 g(stream_seed(seed, 3, b));
        for (uint64_t j = b * BS; j < std::min(nS, (b + 1) * BS); j++) {
            uint64_t rank;
            if (uniform) {
                rank = g.bounded(nR);
            } else {
                uint64_t u = g.next() >> 1;
                uint64_t gi = u >> (63 - glog);
                uint64_t lo = guide[gi], hi = guide[gi + 1];
                while (lo < hi) {   // first r with thr[r] > u
                    uint64_t mid = (lo + hi) >> 1;
                    if (thr[mid] > u) hi = mid; else lo = mid + 1;
                }
                rank = lo;
            }
            S[j].key = R[sigma[rank]].key;
            S[j].rid = (uint32_t)j;
        }
    });
    return 0;
}
