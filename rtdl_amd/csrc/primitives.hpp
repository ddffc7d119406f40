// Device-wide primitives used by the build and emit stages.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace rtdl {

// Stable LSD radix sort of (key, value) pairs on the low `bits` bits of the
// keys (8 bits per pass, LDS-resident per-tile ranking by wave64 ballots).
// `keys`/`values` hold the input and receive the sorted output.
void radix_sort_pairs(uint64_t* keys, uint32_t* values, size_t n, int bits, hipStream_t stream);

// out[i] = sum(in[0..i)) for i in [0, n]; out has n + 1 entries, so out[n]
// is the total.
void exclusive_scan(const uint32_t* in, uint64_t* out, size_t n, hipStream_t stream);

// Fills out[i] = i for i in [0, n).
void iota(uint32_t* out, size_t n, hipStream_t stream);

// Returns true when keys[i] <= keys[i + 1] for all i (device reduction).
bool is_nondecreasing(const uint32_t* keys, size_t n, hipStream_t stream);

}  // namespace rtdl
