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
int check_resolved(const Resolved &r);
int validate_sizes(uint64_t nR, uint64_t nS);
void fill_times(const PhaseTimes &t, spj_phase_times_t *o);
Context &default_context();
std::mutex &default_context_mutex();

}  // namespace spj
