// type_defs.hpp -- the reference's boundary types (type_defs.hpp:1-83 of the
// reference): the tuple kv_pair_t, the operator enum used as phase labels,
// and ResultBuffers_t.  No UPMEM types.
#pragma once

#include <stdint.h>

#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#define NUM_DPU_RANK 64   // DPUs per rank in the reference; only a design input here

// Operator names of the reference's query plans (type_defs.hpp:13-51); the GPU
// phases that replace them are listed in join_internals.hpp.
enum FUNCTION_TYPES {
    DPU_FUNC_LOCAL_HASH_PARTITIONING,
    DPU_FUNC_GLOBAL_HASH_PARTITIONING,
    DPU_FUNC_PHJ_BUILD_HASH_TABLE_LINEAR_PROBE,
    DPU_FUNC_PHJ_PROBE_HASH_TABLE,
    DPU_FUNC_NPHJ_BUILD_HASH_TABLE,
    DPU_FUNC_NPHJ_PROBE_HASH_TABLE,
    DPU_FUNC_PACKETWISE_GLOBAL_HASH_PARTITIONING,
    DPU_FUNC_PACKETWISE_LOCAL_HASH_PARTITIONING,
    DPU_FUNC_PACKETWISE_PHJ_BUILD_HASH_TABLE,
    DPU_FUNC_PHJ_PROBE_HASH_TABLE_INNER_LINEAR_PROBE,
    DPU_FUNC_NESTED_LOOP_JOIN,
    DPU_FUNC_MPSM_JOIN_LOCAL_PARTITIONING,
    DPU_FUNC_MPSM_JOIN_SORT_PROBE,
    DPU_FUNC_GLB_PARTITION_COUNT,
    DPU_FUNC_GLB_PARTITION_PACKET,
    DPU_FUNC_BG_BROADCAST_COUNT,
    DPU_FUNC_BG_BROADCAST_ALIGN,
    DPU_FUNC_GLB_CHIPWISE_PARTITION_COUNT,
    DPU_FUNC_GLB_CHIPWISE_PARTITION_PACKET,
    DPU_FUNC_FINISH_JOIN,
    HOST_FUNC_INVALIDATE_STACKNODE,
    HOST_FUNC_ROTATE_AND_STREAM,
    HOST_FUNC_CALCULATE_PAGE_HISTOGRAM,
    HOST_FUNC_CALCULATE_BANKGROUP_HISTOGRAM,
    HOST_FUNC_LOAD_COLUMN,
    HOST_FUNC_SEND_DATA_OPT,
    HOST_FUNC_RECV_DATA_OPT,
    HOST_FUNC_CHIPWISE_BROADCAST,
    HOST_FUNC_BROADCAST_DATA,
    COMPOUND_FUNC_RNS_JOIN,
    COMPOUND_FUNC_IDP_JOIN,
    COMPOUND_FUNC_GLB_PARTITION,
    COMPOUND_FUNC_GLB_CHIPWISE_PARTITION,
    COMPOUND_FUNC_BG_BROADCAST,
    DPU_FUNC_PACKETWISE_NPHJ_BUILD,
    DPU_FUNC_PACKETWISE_NPHJ_PROBE,
    CONTROL_FUNC_SYNC_THREADS,
};

typedef struct {
    uint32_t lvalue;   // key
    uint32_t rvalue;   // rid (input) / S rid (output pair: lvalue = R rid)
} kv_pair_t;

namespace pidjoin {
// One or more pair buffers and their sizes in BYTES (type_defs.hpp:82).  This
// library returns one buffer whose byte size is exactly 8 * pairs (no
// padding); the caller frees each buffer with free().
typedef std::pair<std::vector<kv_pair_t *>, std::vector<int64_t>> ResultBuffers_t;
}  // namespace pidjoin
