// spj_join_instance.cpp -- pidjoin::JoinInstance on top of the C ABI.
#include <stdio.h>
#include <stdlib.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "pidjoin.hpp"

namespace pidjoin {

namespace {

spj_join_design_t to_c(const join_design_t &d)
{
    spj_join_design_t c;
    c.input_size = d.input_size;
    c.rank_num = d.rank_num;
    c.zipf_factor = d.zipf_factor;
    c.rs_ratio = d.rs_ratio;
    c.rank_set = d.rank_set;
    c.bank_set = d.bank_set;
    c.join_alg = d.join_alg;
    return c;
}

void check(int rc, const char *what)
{
    if (rc != SPJ_OK) throw std::runtime_error(std::string(what) + ": " + spj_last_error());
}

}  // namespace

JoinInstance::JoinInstance() : JoinInstance(16, 1, 16) {}

JoinInstance::JoinInstance(int n) : JoinInstance(n, 1, n) {}

JoinInstance::JoinInstance(int, int) : num_rank_allocated(0), num_rankgroup(0), num_rank_in_rankgroup(0), h_(nullptr)
{
    throw std::logic_error("JoinInstance(int num_rank_allocated, int rank_group_id) is deprecated; use "
                           "JoinInstance(num_rank_allocated, num_rankgroup, num_rank_in_rankgroup)");
}

JoinInstance::JoinInstance(int n, int groups, int per_group)
    : num_rank_allocated(n), num_rankgroup(groups), num_rank_in_rankgroup(per_group), h_(nullptr)
{
    h_ = spj_create(n, groups, per_group);
    if (!h_) throw std::runtime_error(std::string("JoinInstance: ") + spj_last_error());
}

JoinInstance::~JoinInstance() { spj_destroy(h_); }

void JoinInstance::SetRankAllocated(int n) { num_rank_allocated = n; }

void JoinInstance::LoadColumn(void *data_ptr, int64_t num_tuples, std::string col_name)
{
    check(spj_load_column(h_, data_ptr, num_tuples, col_name.c_str()), "LoadColumn");
}

void JoinInstance::PrepareJoin(join_design_t d)
{
    spj_join_design_t c = to_c(d);
    check(spj_prepare(h_, &c), "PrepareJoin");
}

void JoinInstance::StartJoin(join_design_t d)
{
    spj_join_design_t c = to_c(d);
    check(spj_start(h_, &c), "StartJoin");
}

ResultBuffers_t JoinInstance::FinishJoin()
{
    // one exact buffer, or one zero-padded buffer per rank (SetResultLayout)
    const int cap = num_rank_allocated > 1 ? num_rank_allocated : 1;
    std::vector<spj_pair_t *> bufs((size_t)cap, nullptr);
    std::vector<uint64_t> bytes((size_t)cap, 0);
    int n = 0;
    check(spj_finish_buffers(h_, bufs.data(), bytes.data(), cap, &n), "FinishJoin");
    result_bufferpool.first.resize((size_t)n);
    result_bufferpool.second.resize((size_t)n);
    for (int i = 0; i < n; i++) {
        result_bufferpool.first[(size_t)i] = reinterpret_cast<kv_pair_t *>(bufs[(size_t)i]);
        result_bufferpool.second[(size_t)i] = (int64_t)bytes[(size_t)i];
    }
    return result_bufferpool;
}

void JoinInstance::SetResultLayout(int layout) { check(spj_set_result_layout(h_, layout), "SetResultLayout"); }

void JoinInstance::ExecuteJoin(std::string join_type, join_design_t d)
{
    if (join_type != "phj" && join_type != "smj") throw std::runtime_error("Not supported join algorithm " + join_type);
    d.join_alg = (join_type == "smj") ? JOIN_DESIGN_JOIN_ALGORITHM_SMJ : JOIN_DESIGN_JOIN_ALGORITHM_PHJ;
    StartJoin(d);
    FinishJoin();
}

// Timeline of the last join (queryHostFunc.cc:183-241, the stdout branch):
// one "rank_group-0-idx-start-stop-OPERATOR-inputs" line per plan node, times
// in ms from `start`, then the partition transfer lines.
void JoinInstance::RecordLogs(struct timespec start, int rank_group)
{
    uint64_t len = 0;
    if (spj_timeline(h_, (int64_t)start.tv_sec, (int64_t)start.tv_nsec, rank_group, nullptr, 0, &len) != SPJ_OK) return;
    std::string text(len + 1, '\0');
    if (spj_timeline(h_, (int64_t)start.tv_sec, (int64_t)start.tv_nsec, rank_group, &text[0], text.size(), &len) !=
        SPJ_OK)
        return;
    fputs(text.c_str(), stdout);
    fflush(stdout);
}

}  // namespace pidjoin
