// Device-wide primitives: LSD radix sort with LDS tile ranking, multi-level
// exclusive scan, iota and a sortedness check.  All are written for wave64:
// the radix scatter ranks keys inside a wavefront with 8 ballots (one per
// digit bit) and combines the four wavefronts of a 256-thread block in LDS.
#include "common.hpp"
#include "primitives.hpp"

namespace rtdl {
namespace {

constexpr int kSortBlock = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortBlock * kSortItems;
constexpr int kWaves = kSortBlock / 64;

__global__ __launch_bounds__(kSortBlock) void radix_hist_kernel(const uint64_t* __restrict__ keys,
                                                               size_t n, int shift,
                                                               uint32_t* __restrict__ hist,
                                                               uint32_t tiles) {
  __shared__ uint32_t counts[256];
  const uint32_t t = threadIdx.x;
  counts[t] = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * kSortTile;
#pragma unroll 4
  for (int i = 0; i < kSortItems; ++i) {
    size_t idx = base + (size_t)i * kSortBlock + t;
    if (idx < n) atomicAdd(&counts[(keys[idx] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)t * tiles + blockIdx.x] = counts[t];
}

__global__ __launch_bounds__(kSortBlock) void radix_scatter_kernel(
    const uint64_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint64_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, size_t n, int shift,
    const uint64_t* __restrict__ offsets, uint32_t tiles) {
  __shared__ uint64_t base[256];
  __shared__ uint32_t running[256];
  __shared__ uint32_t wave_count[kWaves][256];
  const uint32_t t = threadIdx.x;
  const uint32_t wave = t / 64;
  base[t] = offsets[(size_t)t * tiles + blockIdx.x];
  running[t] = 0;
  const size_t tile_base = (size_t)blockIdx.x * kSortTile;
  const uint64_t lt_mask = __lanemask_lt();
  for (int i = 0; i < kSortItems; ++i) {
#pragma unroll
    for (int w = 0; w < kWaves; ++w) wave_count[w][t] = 0;
    __syncthreads();
    const size_t idx = tile_base + (size_t)i * kSortBlock + t;
    const bool valid = idx < n;
    const uint64_t key = valid ? keys_in[idx] : 0;
    const uint32_t digit = (uint32_t)(key >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (digit >> b) & 1u;
      const uint64_t votes = __ballot(bit);
      peers &= bit ? votes : ~votes;
    }
    const uint32_t rank = __popcll(peers & lt_mask);
    if (valid && rank == 0) wave_count[wave][digit] = __popcll(peers);
    __syncthreads();
    {
      uint32_t s = running[t];
#pragma unroll
      for (int w = 0; w < kWaves; ++w) {
        const uint32_t c = wave_count[w][t];
        wave_count[w][t] = s;
        s += c;
      }
      running[t] = s;
    }
    __syncthreads();
    if (valid) {
      const uint64_t pos = base[digit] + wave_count[wave][digit] + rank;
      keys_out[pos] = key;
      vals_out[pos] = vals_in[idx];
    }
    __syncthreads();
  }
}

// ----------------------------------------------------------------- scan --

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ inline uint64_t wave_inclusive_scan(uint64_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t up = __shfl_up(v, d, 64);
    if (lane >= d) v += up;
  }
  return v;
}

// Block-wide exclusive scan of one value per thread; returns the prefix and
// writes the block total into *total.
__device__ inline uint64_t block_exclusive_scan(uint64_t v, uint64_t* total) {
  __shared__ uint64_t wave_sums[kScanBlock / 64];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x / 64;
  const uint64_t inc = wave_inclusive_scan(v);
  if (lane == 63) wave_sums[wave] = inc;
  __syncthreads();
  uint64_t before = 0;
  uint64_t all = 0;
#pragma unroll
  for (int w = 0; w < kScanBlock / 64; ++w) {
    if (w < wave) before += wave_sums[w];
    all += wave_sums[w];
  }
  *total = all;
  __syncthreads();
  return before + inc - v;
}

template <class Tin>
__global__ __launch_bounds__(kScanBlock) void scan_tile_kernel(const Tin* __restrict__ in,
                                                              uint64_t* __restrict__ out, size_t n,
                                                              uint64_t* __restrict__ tile_sums) {
  const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
  uint64_t vals[kScanItems];
  uint64_t sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const size_t idx = base + i;
    vals[i] = idx < n ? (uint64_t)in[idx] : 0;
    sum += vals[i];
  }
  uint64_t total;
  uint64_t prefix = block_exclusive_scan(sum, &total);
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const size_t idx = base + i;
    if (idx < n) out[idx] = prefix;
    prefix += vals[i];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ void scan_add_kernel(uint64_t* __restrict__ out, size_t n,
                                const uint64_t* __restrict__ tile_offsets, size_t tiles) {
  const size_t idx = (size_t)blockIdx.x * kScanTile + threadIdx.x;
  const uint64_t add = tile_offsets[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const size_t j = idx + (size_t)i * kScanBlock;
    if (j < n) out[j] += add;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = tile_offsets[tiles];
}

__global__ void scan_total_kernel(uint64_t* __restrict__ out, size_t n,
                                  const uint64_t* __restrict__ tile_sums) {
  out[n] = tile_sums[0];
}

template <class Tin>
void scan_impl(const Tin* in, uint64_t* out, size_t n, hipStream_t stream) {
  if (n == 0) {
    RTDL_HIP(hipMemsetAsync(out, 0, sizeof(uint64_t), stream));
    return;
  }
  const size_t tiles = (n + kScanTile - 1) / kScanTile;
  DeviceArray<uint64_t> sums(tiles);
  scan_tile_kernel<Tin><<<(unsigned)tiles, kScanBlock, 0, stream>>>(in, out, n, sums.get());
  RTDL_LAUNCH_CHECK();
  if (tiles == 1) {
    scan_total_kernel<<<1, 1, 0, stream>>>(out, n, sums.get());
    RTDL_LAUNCH_CHECK();
  } else {
    DeviceArray<uint64_t> offs(tiles + 1);
    scan_impl<uint64_t>(sums.get(), offs.get(), tiles, stream);
    scan_add_kernel<<<(unsigned)tiles, kScanBlock, 0, stream>>>(out, n, offs.get(), tiles);
    RTDL_LAUNCH_CHECK();
  }
  // Temporaries are returned to the pool only after the stream has consumed
  // them: the pool hands a block back out only to work ordered after this.
  RTDL_HIP(hipStreamSynchronize(stream));
}

__global__ void scan_zero_kernel(uint64_t* out) { out[0] = 0; }

template <class Tin>
void scan_async_level(const Tin* in, uint64_t* out, size_t n, ScanWorkspace& ws, size_t level,
                      hipStream_t stream) {
  if (n == 0) {
    scan_zero_kernel<<<1, 1, 0, stream>>>(out);
    RTDL_LAUNCH_CHECK();
    return;
  }
  require(level < (size_t)ws.levels, "scan workspace too small");
  const size_t tiles = (n + kScanTile - 1) / kScanTile;
  require(tiles <= ws.tiles[level], "scan workspace too small");
  scan_tile_kernel<Tin><<<(unsigned)tiles, kScanBlock, 0, stream>>>(in, out, n, ws.sums[level]);
  RTDL_LAUNCH_CHECK();
  if (tiles == 1) {
    scan_total_kernel<<<1, 1, 0, stream>>>(out, n, ws.sums[level]);
    RTDL_LAUNCH_CHECK();
  } else {
    scan_async_level<uint64_t>(ws.sums[level], ws.offs[level], tiles, ws, level + 1, stream);
    scan_add_kernel<<<(unsigned)tiles, kScanBlock, 0, stream>>>(out, n, ws.offs[level], tiles);
    RTDL_LAUNCH_CHECK();
  }
}

__global__ void fill_u32_kernel(uint32_t* out, size_t n, uint32_t value) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = value;
}

__global__ void iota_kernel(uint32_t* out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (uint32_t)i;
}

__global__ void nondecreasing_kernel(const uint32_t* keys, size_t n, uint32_t* flag) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 1 < n && keys[i] > keys[i + 1]) *flag = 0u;
}

}  // namespace

void radix_sort_pairs(uint64_t* keys, uint32_t* values, size_t n, int bits, hipStream_t stream) {
  if (n <= 1 || bits <= 0) return;
  const uint32_t tiles = (uint32_t)((n + kSortTile - 1) / kSortTile);
  DeviceArray<uint64_t> keys_tmp(n);
  DeviceArray<uint32_t> vals_tmp(n);
  DeviceArray<uint32_t> hist((size_t)256 * tiles);
  DeviceArray<uint64_t> offs((size_t)256 * tiles + 1);
  uint64_t* kin = keys;
  uint32_t* vin = values;
  uint64_t* kout = keys_tmp.get();
  uint32_t* vout = vals_tmp.get();
  int passes = 0;
  for (int shift = 0; shift < bits; shift += 8, ++passes) {
    radix_hist_kernel<<<tiles, kSortBlock, 0, stream>>>(kin, n, shift, hist.get(), tiles);
    RTDL_LAUNCH_CHECK();
    scan_impl<uint32_t>(hist.get(), offs.get(), (size_t)256 * tiles, stream);
    radix_scatter_kernel<<<tiles, kSortBlock, 0, stream>>>(kin, vin, kout, vout, n, shift,
                                                           offs.get(), tiles);
    RTDL_LAUNCH_CHECK();
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  if (passes & 1) {
    RTDL_HIP(hipMemcpyAsync(keys, kin, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, stream));
    RTDL_HIP(hipMemcpyAsync(values, vin, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream));
  }
  RTDL_HIP(hipStreamSynchronize(stream));
}

void exclusive_scan(const uint32_t* in, uint64_t* out, size_t n, hipStream_t stream) {
  scan_impl<uint32_t>(in, out, n, stream);
}

size_t ScanWorkspace::layout(size_t max_n) {
  levels = 0;
  size_t n = max_n ? max_n : 1;
  size_t total = 0;
  for (;;) {
    const size_t t = (n + kScanTile - 1) / kScanTile;
    require(levels < kMaxLevels, "scan workspace: too many levels");
    tiles[levels++] = t;
    total += 2 * t + 1;
    if (t <= 1) break;
    n = t;
  }
  return total;
}

void ScanWorkspace::point(uint64_t* p) {
  for (int l = 0; l < levels; ++l) {
    sums[l] = p;
    p += tiles[l];
    offs[l] = p;
    p += tiles[l] + 1;
  }
}

void ScanWorkspace::reserve(size_t max_n) {
  if (max_n <= capacity_ && levels) return;
  storage_.resize(layout(max_n));
  point(storage_.get());
  capacity_ = max_n;
}

void ScanWorkspace::carve(Carver& c, size_t max_n) {
  const size_t words = layout(max_n);
  c.take(storage_, words);
  if (c.base) point(storage_.get());
  capacity_ = max_n;
}

void exclusive_scan_async(const uint32_t* in, uint64_t* out, size_t n, ScanWorkspace& ws,
                          hipStream_t stream) {
  require(n <= ws.capacity(), "scan workspace too small");
  scan_async_level<uint32_t>(in, out, n, ws, 0, stream);
}

void SortWorkspace::reserve(size_t max_n) {
  if (max_n <= capacity_ && keys_tmp.size()) return;
  const size_t n = max_n ? max_n : 1;
  const size_t tiles = (n + kSortTile - 1) / kSortTile;
  keys_tmp.resize(n);
  vals_tmp.resize(n);
  hist.resize(256 * tiles);
  offs.resize(256 * tiles + 1);
  scan.reserve(256 * tiles);
  capacity_ = max_n;
}

void SortWorkspace::carve(Carver& c, size_t max_n) {
  const size_t n = max_n ? max_n : 1;
  const size_t tiles = (n + kSortTile - 1) / kSortTile;
  c.take(keys_tmp, n);
  c.take(vals_tmp, n);
  c.take(hist, 256 * tiles);
  c.take(offs, 256 * tiles + 1);
  scan.carve(c, 256 * tiles);
  capacity_ = max_n;
}

void radix_sort_pairs_async(uint64_t* keys, uint32_t* values, size_t n, int bits,
                            SortWorkspace& ws, hipStream_t stream) {
  if (n <= 1 || bits <= 0) return;
  require(n <= ws.capacity(), "sort workspace too small");
  const uint32_t tiles = (uint32_t)((n + kSortTile - 1) / kSortTile);
  int passes = (bits + 7) / 8;
  if (passes & 1) ++passes;  // an extra pass over zero digits is a stable copy
  uint64_t* kin = keys;
  uint32_t* vin = values;
  uint64_t* kout = ws.keys_tmp.get();
  uint32_t* vout = ws.vals_tmp.get();
  for (int pass = 0; pass < passes; ++pass) {
    const int shift = 8 * pass;
    radix_hist_kernel<<<tiles, kSortBlock, 0, stream>>>(kin, n, shift, ws.hist.get(), tiles);
    RTDL_LAUNCH_CHECK();
    exclusive_scan_async(ws.hist.get(), ws.offs.get(), (size_t)256 * tiles, ws.scan, stream);
    radix_scatter_kernel<<<tiles, kSortBlock, 0, stream>>>(kin, vin, kout, vout, n, shift,
                                                           ws.offs.get(), tiles);
    RTDL_LAUNCH_CHECK();
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
}

void fill_u32(uint32_t* out, size_t n, uint32_t value, hipStream_t stream) {
  if (!n) return;
  fill_u32_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(out, n, value);
  RTDL_LAUNCH_CHECK();
}

void iota(uint32_t* out, size_t n, hipStream_t stream) {
  if (!n) return;
  iota_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(out, n);
  RTDL_LAUNCH_CHECK();
}

bool is_nondecreasing(const uint32_t* keys, size_t n, hipStream_t stream) {
  if (n <= 1) return true;
  DeviceArray<uint32_t> flag(1);
  const uint32_t one = 1;
  RTDL_HIP(hipMemcpyAsync(flag.get(), &one, sizeof(one), hipMemcpyHostToDevice, stream));
  nondecreasing_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(keys, n, flag.get());
  RTDL_LAUNCH_CHECK();
  uint32_t result = 0;
  RTDL_HIP(hipMemcpyAsync(&result, flag.get(), sizeof(result), hipMemcpyDeviceToHost, stream));
  RTDL_HIP(hipStreamSynchronize(stream));
  return result != 0;
}

}  // namespace rtdl
