// join_internals.hpp -- pidjoin::JoinInstance, the reference's public join
// API (join_internals.hpp:301-323, src/query.cc), on one MI355X.
//
//   JoinInstance(ranks, rank_groups, ranks_per_group)   query.cc:842
//   JoinInstance(ranks, rank_group_id)  -- deprecated    query.cc:733
//   LoadColumn(ptr, n, "left" | "right")                query.cc:213
//   PrepareJoin(design)                                  query.cc:178
//   StartJoin(design)   -- asynchronous                  query.cc:188
//   FinishJoin() -> ResultBuffers_t                      query.cc:193
//   ExecuteJoin("phj" | "smj", design)                   query.cc:199
//   RecordLogs(start, rank_group)                        queryHostFunc.cc:183
//
// Differences a caller can see: the result is one exact-size buffer of
// (R rid, S rid) pairs (no zero padding, byte size = 8 * pairs); errors
// throw std::runtime_error instead of exit(-1).
#pragma once

#include <time.h>

#include <string>

#include "idpHandler.hpp"
#include "spidjoin.h"
#include "type_defs.hpp"

namespace pidjoin {

class JoinInstance {
public:
    JoinInstance();
    explicit JoinInstance(int num_rank_allocated_);
    // Deprecated in the reference (join_internals.hpp:305), whose body prints
    // an error and exits (query.cc:733-736); here it throws std::logic_error.
    [[deprecated]] JoinInstance(int num_rank_allocated_, int rank_group_id);
    JoinInstance(int num_rank_allocated_, int num_rankgroup, int num_rank_in_rankgroup);
    ~JoinInstance();
    JoinInstance(const JoinInstance &) = delete;
    JoinInstance &operator=(const JoinInstance &) = delete;

    int num_rank_allocated;
    int num_rankgroup;
    int num_rank_in_rankgroup;
    void SetRankAllocated(int num_rank_allocated_);

    void LoadColumn(void *data_ptr, int64_t num_tuples, std::string col_name);
    void ExecuteJoin(std::string join_type, join_design_t join_design);
    void PrepareJoin(join_design_t join_design);
    void StartJoin(join_design_t join_design);
    ResultBuffers_t FinishJoin();
    void RecordLogs(struct timespec start, int rank_group);
    // Not in the reference: FinishJoin's buffers.  SPJ_RESULT_EXACT (default)
    // gives one exact buffer; SPJ_RESULT_RANK_PADDED the reference's layout,
    // one zero-padded buffer per rank (see spidjoin.h).
    void SetResultLayout(int layout);

private:
    spj_instance_t *h_;
    ResultBuffers_t result_bufferpool;
};

}  // namespace pidjoin
