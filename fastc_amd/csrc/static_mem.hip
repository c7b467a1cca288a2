// Device memory the library keeps in the code object instead of allocating:
// the process-wide BC7 watermark counter, the BC7 single-colour endpoint
// table, and the fixed scratch area of fh_compress_device_noalloc.
//
// In host code the names below are host-side stand-ins, not device
// addresses.  static_device_memory() looks each one up with
// hipGetSymbolAddress once per device (with that device current) and every
// other use goes through the pointers it returns.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "internal.h"

namespace fh {
namespace {

// wm_words[0]: the reference's process-global gWMVal (Compressor.cpp:140),
// initially 0xFFFFFFFF; wm_words[1]: staging word for a caller's counter.
__device__ __align__(256) uint32_t wm_words[2] = {0xFFFFFFFFu, 0u};
__device__ __align__(256) uint32_t sc_table[kSingleColorEntries];
// fh_compress_device_noalloc's work area (kNoAllocScratchBytes per device)
__device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

std::mutex g_mu;
StaticDeviceMemory g_mem[kMaxDevices];
bool g_done[kMaxDevices];

}  // namespace

hipError_t static_device_memory(int dev, StaticDeviceMemory *out) {
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_done[dev]) {
    StaticDeviceMemory m;
    hipError_t e = hipGetSymbolAddress(reinterpret_cast<void **>(&m.wm), HIP_SYMBOL(wm_words));
    if (e != hipSuccess) return e;
    e = hipGetSymbolAddress(reinterpret_cast<void **>(&m.sctab), HIP_SYMBOL(sc_table));
    if (e != hipSuccess) return e;
    e = hipGetSymbolAddress(reinterpret_cast<void **>(&m.scratch), HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    m.scratch_bytes = kNoAllocScratchBytes;
    g_mem[dev] = m;
    g_done[dev] = true;
  }
  *out = g_mem[dev];
  return hipSuccess;
}

}  // namespace fh
