// Device-wide primitives used by the build and emit stages.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "common.hpp"

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

// ---- asynchronous variants (no host synchronisation, no allocation, no
// memset nodes), safe to enqueue back to back and to capture in a hipGraph.
// The workspace must stay alive until the stream has executed the work.

// Device memory for exclusive_scan_async over up to `max_n` inputs.
class ScanWorkspace {
 public:
  static constexpr int kMaxLevels = 8;  // 2^64 inputs at 256 per tile and level
  void reserve(size_t max_n);
  // Lays the workspace out in `c` (borrowed storage; see Carver).
  void carve(Carver& c, size_t max_n);
  size_t capacity() const { return capacity_; }
  // per level: tile sums and their exclusive prefix
  uint64_t* sums[kMaxLevels] = {};
  uint64_t* offs[kMaxLevels] = {};
  size_t tiles[kMaxLevels] = {};
  int levels = 0;

 private:
  size_t layout(size_t max_n);  // fills tiles / levels, returns the words needed
  void point(uint64_t* p);      // sets sums / offs into storage at p
  size_t capacity_ = 0;
  DeviceArray<uint64_t> storage_;
};

void exclusive_scan_async(const uint32_t* in, uint64_t* out, size_t n, ScanWorkspace& ws,
                          hipStream_t stream);

// Device memory for radix_sort_pairs_async over up to `max_n` pairs.
class SortWorkspace {
 public:
  void reserve(size_t max_n);
  void carve(Carver& c, size_t max_n);  // borrowed storage laid out in `c`
  size_t capacity() const { return capacity_; }
  DeviceArray<uint64_t> keys_tmp;
  DeviceArray<uint32_t> vals_tmp;
  DeviceArray<uint32_t> hist;
  DeviceArray<uint64_t> offs;
  ScanWorkspace scan;

 private:
  size_t capacity_ = 0;
};

// Sorts on the low `bits` bits; an even number of 8-bit passes is always
// used so that the result lands in keys/values without a copy.
void radix_sort_pairs_async(uint64_t* keys, uint32_t* values, size_t n, int bits,
                            SortWorkspace& ws, hipStream_t stream);

// out[i] = value for i in [0, n) (a fill kernel; graph-safe replacement for
// hipMemsetAsync).
void fill_u32(uint32_t* out, size_t n, uint32_t value, hipStream_t stream);

}  // namespace rtdl
