// Allocation-free simulator entry point: ptyk_simulate with device inputs and outputs and its
// exit-wave staging in a file-scope device array, for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving or data-generation loops).
//
// The array lives in the library's code object: every GPU the library runs on holds its own copy
// (PTYK_NOALLOC_SCRATCH_BYTES) until the process ends.  Host code never passes the array's name to
// a kernel or a copy: its device address is looked up once per device at set-up
// (noalloc_setup, called by ptyk_create with the device current) and that address is used.
// All calls on one GPU share the array, so calls that overlap on one GPU (two streams, two
// threads) race; the contents left by the last call are not cleared (every call overwrites the
// part it uses before reading it).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/ptyk.h"
#include "ptyk_internal.h"

namespace ptyk {

constexpr size_t kScratchBytes = PTYK_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) float2 scratch[kScratchBytes / sizeof(float2)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

}  // namespace

// Look up the scratch array's address on the current device (once per device).
hipError_t noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return hipSuccess;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!p) return hipErrorInvalidSymbol;
  g_ptr[device] = p;
  return hipSuccess;
}

void* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  return g_ptr[device];
}

size_t noalloc_scratch_bytes() { return kScratchBytes; }

}  // namespace ptyk
