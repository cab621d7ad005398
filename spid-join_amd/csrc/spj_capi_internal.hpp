// spj_capi_internal.hpp -- helpers shared by the C ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>

#include "spidjoin.h"
#include "spj_engine.hpp"

namespace spj {

int fail(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);
int err_from_flags(uint32_t flags, size_t inst);
// Parameters for one instance: tuner table / defaults, then explicit fields of cfg.
Resolved resolve(const spj_config_t *cfg, uint64_t nR, uint64_t nS, double skew_hint);
// r's fields are valid; with nR, also the algorithm's size limit for |R| = nR.
int check_resolved(const Resolved &r, uint64_t nR = 0);
// One instance on the default context (held locked by the caller), in one
// radix pass or two (s.cfg.passes).
hipError_t run_one(const InstSpec &s, uint32_t &count, uint32_t &err, PhaseTimes *times);
// AUTO's choice of the no-partition join for a single call (see spj_capi.cpp).
bool auto_no_partition_single(const spj_config_t *cfg, uint64_t nR, uint64_t nS);
int validate_sizes(uint64_t nR, uint64_t nS);
void fill_times(const PhaseTimes &t, spj_phase_times_t *o);
Context &default_context();
std::mutex &default_context_mutex();

}  // namespace spj
