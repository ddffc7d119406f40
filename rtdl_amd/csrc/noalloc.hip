// Device scratch of the allocation-free batch entry point
// (rtdl_mi355x_run_batch_noalloc).
//
// Callers that cannot allocate (code under stream capture, tight serving
// loops, harnesses that forbid hipMalloc) run batches in a fixed engine whose
// every device buffer is carved out of the file-scope array `scratch`: job
// descriptors, work items and their sort keys, queue and job counters, the
// Morton keys and order, the exact fix-up lists and the fixed-radius plans,
// stage and parking space.  Each device holds its own copy of the array from
// the moment the library's code object is loaded on it until the process
// ends.  Host code never uses the array's name as an address: at set-up (the
// entry point's first call on a device, with that device current) its device
// address is looked up once with hipGetSymbolAddress and kept per device, and
// that address is what the engine carves and the kernels receive.
#include <mutex>
#include <new>

#include "common.hpp"
#include "engine.hpp"
#include "noalloc.hpp"

namespace rtdl {
namespace {

constexpr int kMaxDevices = 16;
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr size_t kHostStageBytes = size_t(64) << 10;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct DeviceSlot {
  void* scratch_ptr = nullptr;  // this device's copy of `scratch`
  BatchEngine* engine = nullptr;
  alignas(BatchEngine) unsigned char engine_mem[sizeof(BatchEngine)];
  alignas(256) unsigned char host_stage[kHostStageBytes];
  HostJob jobs[kNoAllocMaxJobs];
};

DeviceSlot g_slots[kMaxDevices];
std::mutex g_mutex;

void set_up(DeviceSlot& slot) {
  size_t dev_bytes = 0, host_bytes = 0;
  BatchEngine::fixed_bytes(kNoAllocMaxJobs, kNoAllocMaxQueries, &dev_bytes, &host_bytes);
  require(dev_bytes <= kScratchBytes && host_bytes <= kHostStageBytes,
          "allocation-free engine: limits exceed the scratch array");
  void* ptr = nullptr;
  RTDL_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  require(ptr != nullptr, "allocation-free engine: scratch symbol has no device address");
  slot.scratch_ptr = ptr;
  EngineStorage st;
  st.device = ptr;
  st.device_bytes = kScratchBytes;
  st.host = slot.host_stage;
  st.host_bytes = kHostStageBytes;
  st.max_jobs = kNoAllocMaxJobs;
  st.max_queries = kNoAllocMaxQueries;
  // the null stream: the call creates no stream; the engine creates no events
  slot.engine = new (slot.engine_mem) BatchEngine(nullptr, st);
}

}  // namespace

NoAllocContext noalloc_context() {
  int dev = 0;
  RTDL_HIP(hipGetDevice(&dev));
  require(dev >= 0 && dev < kMaxDevices, "allocation-free engine: device index above 15");
  DeviceSlot& slot = g_slots[dev];
  if (!slot.engine) set_up(slot);
  return NoAllocContext{slot.engine, slot.jobs};
}

std::mutex& noalloc_mutex() { return g_mutex; }

}  // namespace rtdl
