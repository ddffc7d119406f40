// Allocation-free batch engine, one per device (see noalloc.hip).
#pragma once

#include <cstddef>
#include <mutex>

#include "engine.hpp"

namespace rtdl {

// Limits of one allocation-free batch.
constexpr size_t kNoAllocMaxJobs = 128;
constexpr size_t kNoAllocMaxQueries = 16384;

struct NoAllocContext {
  BatchEngine* engine;  // fixed engine over this device's scratch, on the null stream
  HostJob* jobs;        // kNoAllocMaxJobs job slots (static storage)
};

// The current device's engine, set up on its first use.  Callers hold
// noalloc_mutex() across set-up and the whole batch: the engine and its
// scratch serve one batch at a time.
NoAllocContext noalloc_context();
std::mutex& noalloc_mutex();

}  // namespace rtdl
