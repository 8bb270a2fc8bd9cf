// Device scratch of the allocation-free entry points (adg_sweep_noalloc).
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc) get entry points whose
// staging buffers live in this file-scope array instead of library-allocated
// buffers.  Every GPU the library runs on holds its own copy of the array
// until the process ends; its device address is resolved once per device at
// context creation and only that address is used by host code.  All calls on
// one GPU share the array: calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>

#include "context.hpp"

namespace {

constexpr int kMaxDevices = 64;
// 2^22 doubles (32 MiB) per GPU
static __device__ __align__(256) double scratch[adg::kNoallocScratchDoubles];
double* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

}  // namespace

namespace adg {

void noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) throw std::runtime_error("scratch: device id out of range");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  int cur = -1;
  hip_check(hipGetDevice(&cur), "hipGetDevice");
  if (cur != device) throw std::runtime_error("scratch: the device must be current at set-up");
  void* p = nullptr;
  hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  if (!p) throw std::runtime_error("scratch: no device address");
  g_scratch_ptr[device] = static_cast<double*>(p);
}

double* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) throw std::runtime_error("scratch: device id out of range");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[device]) throw std::runtime_error("scratch: not set up on this device");
  return g_scratch_ptr[device];
}

}  // namespace adg
