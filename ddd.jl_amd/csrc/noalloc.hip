// Fixed device scratch for the entry point that allocates nothing
// (ddd_sigma_tilde_noalloc): a file-scope device array whose address is
// looked up once per device when a context is created on it.  Every call on
// a device shares that device's array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kernels.hpp"

namespace ddd {

static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_addr[kMaxDevices] = {};
hipError_t g_status[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_done[device]) {
    // The caller has made `device` current; each device holds its own copy
    // of the array at its own address.
    void* p = nullptr;
    g_status[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_addr[device] = g_status[device] == hipSuccess ? p : nullptr;
    g_done[device] = true;
  }
  return g_status[device];
}

double* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  return g_done[device] ? static_cast<double*>(g_addr[device]) : nullptr;
}

}  // namespace ddd
