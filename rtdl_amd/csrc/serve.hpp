// Served execution: a request loop over resident scenes with continuous
// batching and memory-budget admission.
#pragma once

#include <cstddef>
#include <cstdint>

#include "../../include/rtdl_mi355x.h"

namespace rtdl {

void serve_replay(void* const* scenes, size_t n_scenes, const RtdlServeRequest* requests,
                  size_t n_requests, const RtdlServeConfig& config, RtdlServeResult* results,
                  RtdlServeStats* stats);

}  // namespace rtdl
