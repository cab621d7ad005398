// spj_plan.hpp -- join plans (the query_tree operator sequences that
// include/execution.hpp defines as join_json and bg_join_json and that
// IDPHandler::SetQueryTree selects, src/pipeline.cc:251-279) and the
// per-operator timeline RecordLogs prints (src/queryHostFunc.cc:183-303).
//
// A plan here is validated and mapped onto the fixed GPU phase sequence; it
// does not schedule kernels operator by operator.  Each operator of the plan
// is tied to the GPU phase that does its work, which gives the timeline its
// per-operator start and stop times.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace spj {

enum class PlanOp {
    Send,                // HOST_FUNC_SEND_DATA_OPT (buffer_name left | right): H2D copy
    Sync,                // CONTROL_FUNC_SYNC_THREADS: stream order, no work
    Partition,           // COMPOUND_FUNC_GLB_PARTITION (inputs[1] left | right)
    ChipwisePartition,   // COMPOUND_FUNC_GLB_CHIPWISE_PARTITION (right; heavy S tuples joined in place)
    BgBroadcast,         // COMPOUND_FUNC_BG_BROADCAST (left): heavy-key detection
    ChipwiseBroadcast,   // HOST_FUNC_CHIPWISE_BROADCAST: heavy R tuples to every workgroup (in the kernels)
    RotateAndStream,     // HOST_FUNC_ROTATE_AND_STREAM: partition hand-over (in the scatter)
    RnsJoin,             // COMPOUND_FUNC_RNS_JOIN
    IdpJoin,             // COMPOUND_FUNC_IDP_JOIN
    Recv,                // HOST_FUNC_RECV_DATA_OPT: D2H copy of the result
};

struct PlanNode {
    PlanOp op;
    std::string name;     // operator name as written in the plan
    std::string inputs;   // "input", or "inputs" joined by ',' (as the timeline prints them)
    int side = -1;        // 0 left (R), 1 right (S), -1 neither
};

struct Plan {
    std::vector<PlanNode> nodes;
    bool replicate = false;   // has BG_BROADCAST: heavy-key handling on
};

// Parses and validates a plan document ({"query_tree": [ {"operator": ...}, ... ]}).
// Returns SPJ_OK, or an SPJ_E* code with the message set.
int parse_plan(const char *json, Plan *out);

// The two plans SetQueryTree chooses between (bank_set > 1: replicate).
Plan default_plan(bool replicate);

// Per-node interval in ms from the join's start (CLOCK_MONOTONIC ns t0).
struct NodeTime {
    double start_ms = 0, stop_ms = 0;
};

// The timeline text: one line per node,
// "rank_group-0-index-start-stop-OPERATOR-inputs" with start and stop in ms
// relative to start_ns, then one "rnc_xfer_bytes B rnc_num_packets P" line
// per partitioned relation (B bytes moved as P 8-byte tuples).
std::string timeline_text(const Plan &plan, const std::vector<NodeTime> &t, int64_t t0_ns, int64_t start_ns,
                          int rank_group, const std::vector<uint64_t> &parted);

}  // namespace spj
