// idpHandler.hpp -- only the run configuration of the reference's
// idpHandler.hpp: pidjoin::join_design_t (131-145, inside namespace pidjoin
// as there, 28) and the algorithm ids (126-129).
// The UPMEM device handle, MRAM allocator and argument handlers have no GPU
// counterpart.
#pragma once

#include <stdint.h>

#include <string>

#define JOIN_DESIGN_JOIN_ALGORITHM_PHJ 0
#define JOIN_DESIGN_JOIN_ALGORITHM_NPHJ 1
#define JOIN_DESIGN_JOIN_ALGORITHM_SMJ 2
#define JOIN_DESIGN_JOIN_ALGORITHM_NL 3

namespace pidjoin {

typedef struct join_design_t {
    int64_t input_size = 500000;
    int rank_num = 16;
    int zipf_factor = 0;
    int rs_ratio = 8;
    int rank_set = 1;    // > 1: skew-resistant path (heavy-key handling)
    int bank_set = 1;    // > 1: skew-resistant path (heavy-key handling)
    bool tpch_example = false;
    int join_alg = 0;
    std::string primary_key;
    std::string foreign_key;
    int identical_value_percent = 0;
} join_design_t;

}  // namespace pidjoin
