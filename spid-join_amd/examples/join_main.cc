// join_main.cc -- the reference's join test program's flow (test/join/main.cc:16-116)
// against this library's pidjoin::JoinInstance: same CLI flags
//   -s input_size  -r ranks  -z zipf x100  -t R:S ratio  -R rank_set  -B bank_set
// plus -a join_alg.  Unlike the reference's program it draws S keys Zipf over R
// (the reference leaves Zipf as a stub, main.cc:45-48) and validates the
// result: exact pair count and, for every pair, R.key[r_rid] == S.key[s_rid].
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include <vector>

#include "pidjoin.hpp"

using namespace pidjoin;

int main(int argc, char **argv)
{
    join_design_t d;
    d.input_size = 1 << 20;
    d.rs_ratio = 2;
    int opt;
    while ((opt = getopt(argc, argv, "s:r:z:t:R:B:a:")) != -1) {
        switch (opt) {
        case 's': d.input_size = atoll(optarg); break;
        case 'r': d.rank_num = atoi(optarg); break;
        case 'z': d.zipf_factor = atoi(optarg); break;
        case 't': d.rs_ratio = atoi(optarg); break;
        case 'R': d.rank_set = atoi(optarg); break;
        case 'B': d.bank_set = atoi(optarg); break;
        case 'a': d.join_alg = atoi(optarg); break;
        default: return 2;
        }
    }
    if (spj_init(0) != SPJ_OK) {
        fprintf(stderr, "spj_init: %s\n", spj_last_error());
        return 1;
    }
    const int64_t nR = d.input_size, nS = nR * d.rs_ratio;
    std::vector<kv_pair_t> r((size_t)nR), s((size_t)nS);
    spj_gen_r(reinterpret_cast<spj_tuple_t *>(r.data()), (uint64_t)nR, 1);
    spj_gen_s(reinterpret_cast<spj_tuple_t *>(s.data()), (uint64_t)nS, reinterpret_cast<spj_tuple_t *>(r.data()),
              (uint64_t)nR, d.zipf_factor / 100.0, 2, 0);

    struct timespec start;
    JoinInstance join_instance(d.rank_num, d.rank_set, d.rank_num / d.rank_set);
    join_instance.LoadColumn((void *)r.data(), nR, "left");
    join_instance.LoadColumn((void *)s.data(), nS, "right");
    join_instance.PrepareJoin(d);
    clock_gettime(CLOCK_MONOTONIC, &start);
    join_instance.StartJoin(d);
    ResultBuffers_t res = join_instance.FinishJoin();
    join_instance.RecordLogs(start, 0);

    // validation: the reference's program never reads its result (main.cc:63)
    std::vector<uint32_t> key_of_rid((size_t)nR);
    for (int64_t i = 0; i < nR; i++) key_of_rid[r[(size_t)i].rvalue] = r[(size_t)i].lvalue;
    int64_t pairs = 0, bad = 0;
    for (size_t b = 0; b < res.first.size(); b++) {
        const int64_t n = res.second[b] / (int64_t)sizeof(kv_pair_t);
        for (int64_t i = 0; i < n; i++) {
            const kv_pair_t p = res.first[b][i];
            if (p.lvalue >= (uint64_t)nR || p.rvalue >= (uint64_t)nS || key_of_rid[p.lvalue] != s[p.rvalue].lvalue) bad++;
        }
        pairs += n;
        free(res.first[b]);
    }
    printf("pairs %lld expected %lld bad %lld\n", (long long)pairs, (long long)nS, (long long)bad);
    return (pairs == nS && bad == 0) ? 0 : 1;
}
