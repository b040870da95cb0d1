// noalloc.hip -- the static device scratch of the no-allocation entry point.
//
// appt_spgemm_compute_noalloc takes every scratch buffer of the pipeline from
// one file-scope device array instead of the engine's pool, so the call never
// allocates device memory.  Each device that loads the library holds its own
// copy of the array until the process ends.  Host code must not use the
// array's name as a device address: its address on a device comes from
// hipGetSymbolAddress, taken once per device (with that device current) when
// an engine is created there, and kept in g_addr.
#include <hip/hip_runtime.h>

#include <mutex>

#include "noalloc.hpp"

namespace appt {

static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_addr[kMaxDevices] = {};
}  // namespace

hipError_t static_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_addr[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_addr[device] = p;
  return hipSuccess;
}

void* static_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  return g_addr[device];
}

}  // namespace appt
