// spj_tuner.hpp -- per size / skew bucket parameter table (the autotuner).
#pragma once
#include <stdint.h>

#include "spj_engine.hpp"

namespace spj {

// Parameters for an instance: the tuned entry of its bucket when one exists,
// else default_params().  skew_hint < 0 means unknown.
Resolved tuned_params(uint32_t nR, uint32_t nS, double skew_hint);
// Whether the table holds an entry for the bucket of (nR, nS, skew_hint).
bool tuner_has_entry(uint32_t nR, uint32_t nS, double skew_hint);

}  // namespace spj
