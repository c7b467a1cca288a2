// Per-kernel timing for benchmarks (fh_set_kernel_timing / fh_kernel_time).
//
// While enabled, the launchers bracket each kernel with a pair of HIP events
// recorded on the kernel's own stream; fh_kernel_time waits for the recorded
// events and sums the elapsed times per kernel.  Disabled (the default), the
// launch path only reads one flag.
#include <hip/hip_runtime.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "../../include/fasthip.h"
#include "internal.h"

namespace fh {
namespace {

const char *const kTagNames[kNumKernelTags] = {
    "bc7_scan", "bc7_select", "bc7_setup", "bc7_anneal", "bc7_pack", "bc7_guard",
    "dxt",      "etc1",       "bc7_labels", "bc7_rank", "decode",    "block_error"};

struct Pending {
  int tag;
  hipEvent_t start, stop;
};

std::mutex g_mu;
bool g_enabled = false;
std::vector<Pending> g_pending;
std::vector<hipEvent_t> g_starts;  // start events handed out, by slot
double g_total_ms[kNumKernelTags];
uint32_t g_launches[kNumKernelTags];

void drain_locked() {
  for (const Pending &p : g_pending) {
    float ms = 0.0f;
    if (hipEventSynchronize(p.stop) == hipSuccess &&
        hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
      g_total_ms[p.tag] += ms;
      g_launches[p.tag]++;
    }
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
  g_pending.clear();
}

}  // namespace

int timing_begin(hipStream_t stream) {
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_enabled) return -1;
  // no event nodes inside a graph being captured
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
    return -1;
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return -1;
  if (hipEventRecord(e, stream) != hipSuccess) {
    (void)hipEventDestroy(e);
    return -1;
  }
  g_starts.push_back(e);
  return (int)g_starts.size() - 1;
}

void timing_end(int slot, int tag, hipStream_t stream) {
  if (slot < 0) return;
  std::lock_guard<std::mutex> g(g_mu);
  if (slot >= (int)g_starts.size() || !g_starts[slot]) return;
  hipEvent_t start = g_starts[slot];
  g_starts[slot] = nullptr;
  hipEvent_t stop;
  if (hipEventCreate(&stop) != hipSuccess) {
    (void)hipEventDestroy(start);
    return;
  }
  if (hipEventRecord(stop, stream) != hipSuccess) {
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    return;
  }
  g_pending.push_back({tag, start, stop});
}

}  // namespace fh

extern "C" {

fh_status fh_set_kernel_timing(int enable) {
  std::lock_guard<std::mutex> g(fh::g_mu);
  fh::drain_locked();
  fh::g_starts.clear();
  for (int i = 0; i < fh::kNumKernelTags; i++) {
    fh::g_total_ms[i] = 0.0;
    fh::g_launches[i] = 0;
  }
  fh::g_enabled = enable != 0;
  return FH_OK;
}

fh_status fh_kernel_time(const char *kernel, double *total_ms, uint32_t *launches) {
  if (!kernel || !total_ms || !launches) return FH_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> g(fh::g_mu);
  fh::drain_locked();
  for (int i = 0; i < fh::kNumKernelTags; i++) {
    if (strcmp(kernel, fh::kTagNames[i]) == 0) {
      *total_ms = fh::g_total_ms[i];
      *launches = fh::g_launches[i];
      return FH_OK;
    }
  }
  return FH_ERR_INVALID_ARGUMENT;
}

}  // extern "C"
