// join_main.cc -- the reference's join test program's flow (test/join/main.cc:16-116)
// against this library's pidjoin::JoinInstance: same CLI flags
//   -s input_size  -r ranks  -z zipf x100  -t R:S ratio  -R rank_set  -B bank_set
// plus -a join_alg and -d (call the deprecated two-argument constructor, which
// must throw).  Unlike the reference's program it draws S keys Zipf over R
// (the reference leaves Zipf as a stub, main.cc:45-48) and validates the
// result: the exact pair count, every S rid exactly once, and for every
// pair R.key[r_rid] == S.key[s_rid].
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include <stdexcept>
#include <vector>

#include "pidjoin.hpp"

int main(int argc, char **argv)
{
    pidjoin::join_design_t d;
    d.input_size = 1 << 20;
    d.rs_ratio = 2;
    bool deprecated = false;
    int opt;
    while ((opt = getopt(argc, argv, "s:r:z:t:R:B:a:d")) != -1) {
        switch (opt) {
        case 's': d.input_size = atoll(optarg); break;
        case 'r': d.rank_num = atoi(optarg); break;
        case 'z': d.zipf_factor = atoi(optarg); break;
        case 't': d.rs_ratio = atoi(optarg); break;
        case 'R': d.rank_set = atoi(optarg); break;
        case 'B': d.bank_set = atoi(optarg); break;
        case 'a': d.join_alg = atoi(optarg); break;
        case 'd': deprecated = true; break;
        default: return 2;
        }
    }
    if (spj_init(0) != SPJ_OK) {
        fprintf(stderr, "spj_init: %s\n", spj_last_error());
        return 1;
    }
    if (deprecated) {
        try {
#pragma GCC diagnostic push
#pragma GCC diagnostic ignored "-Wdeprecated-declarations"
            pidjoin::JoinInstance old(d.rank_num, 0);
#pragma GCC diagnostic pop
        } catch (const std::logic_error &e) {
            printf("deprecated constructor refused: %s\n", e.what());
            return 0;
        }
        return 1;
    }
    const int64_t nR = d.input_size, nS = nR * d.rs_ratio;
    std::vector<kv_pair_t> r((size_t)nR), s((size_t)nS);
    spj_gen_r(reinterpret_cast<spj_tuple_t *>(r.data()), (uint64_t)nR, 1);
    spj_gen_s(reinterpret_cast<spj_tuple_t *>(s.data()), (uint64_t)nS, reinterpret_cast<spj_tuple_t *>(r.data()),
              (uint64_t)nR, d.zipf_factor / 100.0, 2, 0);

    struct timespec start;
    pidjoin::JoinInstance join_instance(d.rank_num, d.rank_set, d.rank_num / d.rank_set);
    join_instance.LoadColumn((void *)r.data(), nR, "left");
    join_instance.LoadColumn((void *)s.data(), nS, "right");
    join_instance.PrepareJoin(d);
    clock_gettime(CLOCK_MONOTONIC, &start);
    join_instance.StartJoin(d);
    pidjoin::ResultBuffers_t res = join_instance.FinishJoin();
    join_instance.RecordLogs(start, 0);

    // validation (the reference's program never reads its result, main.cc:63):
    // rids are row indices here, and every S key occurs in R, so the result
    // must hold each S rid exactly once, paired with the R rid of its key
    std::vector<uint32_t> key_of_rid((size_t)nR);
    for (int64_t i = 0; i < nR; i++) key_of_rid[r[(size_t)i].rvalue] = r[(size_t)i].lvalue;
    std::vector<uint8_t> seen((size_t)nS, 0);
    int64_t pairs = 0, bad = 0, repeated = 0;
    for (size_t b = 0; b < res.first.size(); b++) {
        const int64_t n = res.second[b] / (int64_t)sizeof(kv_pair_t);
        for (int64_t i = 0; i < n; i++) {
            const kv_pair_t p = res.first[b][i];
            if (p.lvalue >= (uint64_t)nR || p.rvalue >= (uint64_t)nS || key_of_rid[p.lvalue] != s[p.rvalue].lvalue) {
                bad++;
                continue;
            }
            if (seen[p.rvalue]++) repeated++;
        }
        pairs += n;
        free(res.first[b]);
    }
    int64_t missing = 0;
    for (int64_t j = 0; j < nS; j++) missing += seen[(size_t)j] == 0;
    printf("pairs %lld expected %lld bad %lld repeated %lld missing %lld\n", (long long)pairs, (long long)nS,
           (long long)bad, (long long)repeated, (long long)missing);
    return (pairs == nS && bad == 0 && repeated == 0 && missing == 0) ? 0 : 1;
}
