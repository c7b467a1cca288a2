// Batched encode of many surfaces (textures and their mip levels) of mixed
// formats (include/fasthip.h, fh_batch_*).  The GPU form of FasTC's atomic
// job-list mode (CompressionJobList / CompressAtomic, Compressor.cpp:1542-1574):
// all blocks of a format group are numbered globally and persistent kernels
// take them from device counters, whatever surface they belong to.
//
// A plan owns every device buffer the batch needs, so running it allocates
// nothing, reads nothing back and can be captured into a HIP graph, which the
// plan instantiates once and replays.  The autotuner picks the launch shapes
// (persistent workgroups per CU, annealing steps between refills, workgroup
// size of the DXT/ETC1 kernels) by timing the plan itself.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "../../include/fasthip.h"
#include "internal.h"

namespace fh {
namespace {

struct Group {
  uint32_t total = 0;
  SurfDesc *surfs = nullptr;
  uint32_t *block_surf = nullptr;
};

fh_status bfail(fh_status s, const char *msg) {
  set_error(msg);
  return s;
}

fh_status bfail_hip(hipError_t e, const char *what) {
  char buf[256];
  snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
  set_error(buf);
  return FH_ERR_HIP;
}

size_t block_bytes_of(uint32_t fmt) {
  return fmt == FH_FORMAT_DXT1 || fmt == FH_FORMAT_ETC1 ? 8 : 16;
}

}  // namespace
}  // namespace fh

struct fh_batch {
  int device = 0;
  fh_batch_options opt;
  fh::Group groups[4];  // by format: DXT1, DXT5, ETC1, BPTC
  uint32_t *counters = nullptr;  // fetch counters of the DXT1 / DXT5 / ETC1 kernels
  // BC7 group, processed in calls of at most kBc7MaxBlocksPerCall blocks
  uint32_t bc7_chunk = 0;
  uint32_t *bc7_blocks = nullptr;  // chunk x 16 pixels
  uint8_t *bc7_enc = nullptr;      // chunk x 16 bytes
  uint8_t *bc7_scan = nullptr;     // watermark scan scratch
  void *bc7_work = nullptr;
  uint32_t *wm_counter = nullptr;  // the library's process-wide watermark counter
  fh::LaunchTune tune;
  hipStream_t capture_stream = nullptr;
  hipGraphExec_t graph = nullptr;
  std::vector<void *> allocations;
  std::mutex mu;
};

namespace fh {
namespace {

hipError_t alloc(fh_batch *b, void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes ? bytes : 16);
  if (e == hipSuccess) b->allocations.push_back(*p);
  return e;
}

void release(fh_batch *b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
  for (void *p : b->allocations) (void)hipFree(p);
  b->allocations.clear();
  b->graph = nullptr;
  b->capture_stream = nullptr;
}

// The launch sequence of one run: no allocation, no host synchronisation,
// no memset nodes (counters are cleared by a kernel), so it can be captured.
hipError_t enqueue(fh_batch *b, hipStream_t s) {
  hipError_t e = fill_u32(b->counters, 4, 0u, s);
  if (e != hipSuccess) return e;
  const Group &g1 = b->groups[0], &g5 = b->groups[1], &ge = b->groups[2], &g7 = b->groups[3];
  e = dxt_encode_batch(g1.surfs, g1.block_surf, g1.total, 0, b->counters + 0, b->tune, s);
  if (e == hipSuccess)
    e = dxt_encode_batch(g5.surfs, g5.block_surf, g5.total, 1, b->counters + 1, b->tune, s);
  if (e == hipSuccess)
    e = etc1_encode_batch(ge.surfs, ge.block_surf, ge.total, b->counters + 2, b->tune, s);
  if (e != hipSuccess) return e;
  for (uint32_t g0 = 0; g0 < g7.total; g0 += b->bc7_chunk) {
    const uint32_t n = g7.total - g0 < b->bc7_chunk ? g7.total - g0 : b->bc7_chunk;
    e = bc7_gather(g7.surfs, g7.block_surf, g0, n, b->bc7_blocks, s);
    if (e != hipSuccess) return e;
    BlockRange r;
    r.img = b->bc7_blocks;
    r.out = b->bc7_enc;
    r.width = 4;
    r.blocks_wide = 1;
    r.first_block = 0;
    r.num_blocks = n;
    Bc7Params p;
    p.quality = b->opt.quality;
    p.block_modes = b->opt.block_modes ? (b->opt.block_modes & 0xFFu) : 0xFFu;
    p.seed = b->opt.seed;
    p.block_id_base = g0;
    p.q0_guard = (b->opt.flags & FH_FLAG_NO_Q0_GUARD) ? 0u : 1u;
    p.ranker = (b->opt.flags & FH_FLAG_LEARNED) ? &b->opt.ranker->net : nullptr;
    p.topk = b->opt.ranker_topk ? b->opt.ranker_topk : 4;
    p.threshold = b->opt.ranker_threshold;
    p.wg_per_cu = b->tune.wg_per_cu;
    p.anneal_unroll = b->tune.anneal_unroll;
    p.stats = nullptr;
    p.selections = nullptr;
    // kErrorMetrics (Compressor.cpp:205-208)
    const bool nu = b->opt.error_metric == 1;
    p.nonuniform = nu ? 1u : 0u;
    p.work_counts = nullptr;
    p.metric[0] = nu ? sqrtf(0.3f) : 1.0f;
    p.metric[1] = nu ? sqrtf(0.56f) : 1.0f;
    p.metric[2] = nu ? sqrtf(0.11f) : 1.0f;
    p.metric[3] = 1.0f;
    Bc7Scratch sc;
    const size_t prefix_bytes = ((size_t)b->bc7_chunk + 255) & ~(size_t)255;
    sc.local_prefix = b->bc7_scan;
    sc.chunk_prefix = reinterpret_cast<uint32_t *>(b->bc7_scan + prefix_bytes);
    sc.counter = b->wm_counter;
    sc.num_chunks = (n + 255) / 256;
    e = bc7_encode(r, p, sc, b->bc7_work, s);
    if (e != hipSuccess) return e;
    e = bc7_scatter(g7.surfs, g7.block_surf, g0, n, b->bc7_enc, s);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

}  // namespace
}  // namespace fh

using namespace fh;

extern "C" {

fh_status fh_batch_create(const fh_surface *surfaces, uint32_t count,
                          const fh_batch_options *options, fh_batch **out) {
  if (!out || (!surfaces && count)) return bfail(FH_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  fh_batch_options opt = {};
  if (options) opt = *options;
  if (opt.error_metric > 1) return bfail(FH_ERR_INVALID_ARGUMENT, "unknown error metric");
  if (opt.error_metric == 1 && (opt.flags & FH_FLAG_LEARNED))
    return bfail(FH_ERR_UNSUPPORTED,
                 "the learned search ranks for the uniform metric: the nonuniform metric "
                 "runs the reference search only");
  if ((opt.flags & FH_FLAG_LEARNED) && !opt.ranker)
    return bfail(FH_ERR_INVALID_ARGUMENT, "FH_FLAG_LEARNED needs a ranker");
  if (opt.flags & FH_FLAG_CALLER_WATERMARK)
    return bfail(FH_ERR_UNSUPPORTED, "a batch's own watermark counter is a device word "
                                     "(fh_batch_options.wm_counter)");
  int ndev = 0, dev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return bfail(FH_ERR_NO_DEVICE, "no HIP device available");
  if (hipGetDevice(&dev) != hipSuccess) return bfail(FH_ERR_NO_DEVICE, "hipGetDevice failed");
  if ((opt.flags & FH_FLAG_LEARNED) && opt.ranker->net.device != dev)
    return bfail(FH_ERR_INVALID_ARGUMENT, "ranker belongs to another device");
  // group the surfaces by format, in submission order
  static const int kGroupOf[4] = {0, 1, 2, 3};  // FH_FORMAT_DXT1, DXT5, ETC1, BPTC
  std::vector<SurfDesc> hs[4];
  std::vector<uint32_t> hb[4];
  for (uint32_t i = 0; i < count; i++) {
    const fh_surface &f = surfaces[i];
    if (f.format > FH_FORMAT_BPTC) return bfail(FH_ERR_UNKNOWN_FORMAT, "Unknown compression format");
    if (!f.in || !f.out || f.width == 0 || f.height == 0)
      return bfail(FH_ERR_NO_DATA, "No data sent to compress!");
    if (f.width % 4 || f.height % 4)
      return bfail(FH_ERR_BLOCK_DIMS, "surface width or height is not a multiple of 4");
    const int gi = kGroupOf[f.format];
    SurfDesc d;
    d.img = reinterpret_cast<const uint32_t *>(f.in);
    d.out = f.out;
    d.width = f.width;
    d.blocks_wide = f.width / 4;
    d.first = (uint32_t)hb[gi].size();
    d.nblocks = (f.width / 4) * (f.height / 4);
    if ((uint64_t)d.first + d.nblocks > 0xFFFFFFFFull)
      return bfail(FH_ERR_INVALID_ARGUMENT, "too many blocks in one format group");
    const uint32_t sid = (uint32_t)hs[gi].size();
    hs[gi].push_back(d);
    hb[gi].insert(hb[gi].end(), d.nblocks, sid);
  }
  fh_batch *b = new fh_batch();
  b->device = dev;
  b->opt = opt;
  hipError_t e = hipSuccess;
  for (int gi = 0; gi < 4 && e == hipSuccess; gi++) {
    Group &g = b->groups[gi];
    g.total = (uint32_t)hb[gi].size();
    e = alloc(b, reinterpret_cast<void **>(&g.surfs), hs[gi].size() * sizeof(SurfDesc));
    if (e == hipSuccess) e = alloc(b, reinterpret_cast<void **>(&g.block_surf), hb[gi].size() * 4);
    if (e == hipSuccess && !hs[gi].empty())
      e = hipMemcpy(g.surfs, hs[gi].data(), hs[gi].size() * sizeof(SurfDesc),
                    hipMemcpyHostToDevice);
    if (e == hipSuccess && !hb[gi].empty())
      e = hipMemcpy(g.block_surf, hb[gi].data(), hb[gi].size() * 4, hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) e = alloc(b, reinterpret_cast<void **>(&b->counters), 16);
  const uint32_t n7 = b->groups[3].total;
  b->bc7_chunk = n7 < kBc7MaxBlocksPerCall ? (n7 ? n7 : 1) : kBc7MaxBlocksPerCall;
  const bool learned = (opt.flags & FH_FLAG_LEARNED) != 0;
  if (e == hipSuccess && n7) {
    e = alloc(b, reinterpret_cast<void **>(&b->bc7_blocks), (size_t)b->bc7_chunk * 64);
    if (e == hipSuccess) e = alloc(b, reinterpret_cast<void **>(&b->bc7_enc), (size_t)b->bc7_chunk * 16);
    const size_t prefix_bytes = ((size_t)b->bc7_chunk + 255) & ~(size_t)255;
    if (e == hipSuccess)
      e = alloc(b, reinterpret_cast<void **>(&b->bc7_scan),
                prefix_bytes + ((size_t)b->bc7_chunk + 255) / 256 * 4 + 256);
    if (e == hipSuccess)
      e = alloc(b, &b->bc7_work,
                bc7_work_bytes(b->bc7_chunk, learned, opt.error_metric == 1) + 256);
  }
  if (e == hipSuccess) e = dxt_prepare();
  if (e == hipSuccess) e = etc1_prepare();
  if (e == hipSuccess) e = bc7_prepare();
  if (e == hipSuccess) b->wm_counter = opt.wm_counter ? opt.wm_counter : device_wm_counter(dev);
  if (e == hipSuccess && !b->wm_counter) e = hipErrorNotInitialized;
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    release(b);
    delete b;
    return bfail_hip(e, "batch set-up");
  }
  *out = b;
  return FH_OK;
}

fh_status fh_batch_destroy(fh_batch *b) {
  if (!b) return FH_OK;
  (void)hipDeviceSynchronize();
  release(b);
  delete b;
  return FH_OK;
}

fh_status fh_batch_run(fh_batch *b, int use_graph, void *hip_stream) {
  if (!b) return bfail(FH_ERR_INVALID_ARGUMENT, "null batch");
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != b->device)
    return bfail(FH_ERR_INVALID_ARGUMENT, "batch belongs to another device");
  std::lock_guard<std::mutex> g(b->mu);
  hipStream_t s = reinterpret_cast<hipStream_t>(hip_stream);
  hipError_t e;
  if (!use_graph) {
    e = enqueue(b, s);
    if (e != hipSuccess) return bfail_hip(e, "batch launch");
    return FH_OK;
  }
  if (!b->graph) {
    hipGraph_t graph = nullptr;
    e = hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return bfail_hip(e, "graph capture");
    hipError_t e2 = enqueue(b, b->capture_stream);
    e = hipStreamEndCapture(b->capture_stream, &graph);
    if (e2 != hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&b->graph, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      b->graph = nullptr;
      return bfail_hip(e, "graph capture");
    }
  }
  e = hipGraphLaunch(b->graph, s);
  if (e != hipSuccess) return bfail_hip(e, "graph launch");
  return FH_OK;
}

fh_status fh_batch_autotune(fh_batch *b, void *hip_stream, fh_batch_tune *report) {
  if (!b) return bfail(FH_ERR_INVALID_ARGUMENT, "null batch");
  std::lock_guard<std::mutex> g(b->mu);
  hipStream_t s = reinterpret_cast<hipStream_t>(hip_stream);
  hipEvent_t t0, t1;
  hipError_t e = hipEventCreate(&t0);
  if (e == hipSuccess) e = hipEventCreate(&t1);
  if (e != hipSuccess) return bfail_hip(e, "autotune events");
  auto time_run = [&](float *ms) {
    hipError_t r = enqueue(b, s);  // warm
    if (r == hipSuccess) r = hipEventRecord(t0, s);
    if (r == hipSuccess) r = enqueue(b, s);
    if (r == hipSuccess) r = hipEventRecord(t1, s);
    if (r == hipSuccess) r = hipEventSynchronize(t1);
    if (r == hipSuccess) r = hipEventElapsedTime(ms, t0, t1);
    return r;
  };
  // coordinate descent, one parameter at a time, from the defaults
  static const uint32_t kWg[3] = {8, 16, 32};
  static const uint32_t kUnroll[3] = {1, 2, 4};
  static const uint32_t kSimple[4] = {64, 128, 256, 512};
  LaunchTune best = LaunchTune();
  b->tune = best;
  float best_ms = 0.0f;
  uint32_t trials = 0;
  e = time_run(&best_ms);
  trials++;
  for (int param = 0; param < 3 && e == hipSuccess; param++) {
    const uint32_t *vals = param == 0 ? kWg : (param == 1 ? kUnroll : kSimple);
    const int nv = param == 2 ? 4 : 3;
    for (int i = 0; i < nv && e == hipSuccess; i++) {
      LaunchTune t = best;
      if (param == 0) t.wg_per_cu = vals[i];
      else if (param == 1) t.anneal_unroll = vals[i];
      else t.simple_wg = vals[i];
      if (t.wg_per_cu == best.wg_per_cu && t.anneal_unroll == best.anneal_unroll &&
          t.simple_wg == best.simple_wg)
        continue;
      b->tune = t;
      float ms = 0.0f;
      e = time_run(&ms);
      trials++;
      if (e == hipSuccess && ms < best_ms) {
        best_ms = ms;
        best = t;
      }
    }
  }
  b->tune = best;
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (e != hipSuccess) return bfail_hip(e, "autotune run");
  if (b->graph) {  // re-capture with the chosen shapes
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  if (report) {
    report->wg_per_cu = best.wg_per_cu;
    report->anneal_unroll = best.anneal_unroll;
    report->simple_wg = best.simple_wg;
    report->best_ms = best_ms;
    report->trials = trials;
  }
  return FH_OK;
}

fh_status fh_batch_set_tune(fh_batch *b, const fh_batch_tune *tune) {
  if (!b || !tune) return bfail(FH_ERR_INVALID_ARGUMENT, "null argument");
  if ((tune->wg_per_cu != 8 && tune->wg_per_cu != 16 && tune->wg_per_cu != 32) ||
      (tune->anneal_unroll != 1 && tune->anneal_unroll != 2 && tune->anneal_unroll != 4) ||
      (tune->simple_wg != 64 && tune->simple_wg != 128 && tune->simple_wg != 256 &&
       tune->simple_wg != 512))
    return bfail(FH_ERR_INVALID_ARGUMENT, "launch shape outside the tuned set");
  std::lock_guard<std::mutex> g(b->mu);
  b->tune.wg_per_cu = tune->wg_per_cu;
  b->tune.anneal_unroll = tune->anneal_unroll;
  b->tune.simple_wg = tune->simple_wg;
  if (b->graph) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  return FH_OK;
}

fh_status fh_batch_blocks(const fh_batch *b, uint32_t format, uint32_t *blocks) {
  if (!b || !blocks) return bfail(FH_ERR_INVALID_ARGUMENT, "null argument");
  if (format > FH_FORMAT_BPTC) return bfail(FH_ERR_UNKNOWN_FORMAT, "Unknown compression format");
  *blocks = b->groups[format].total;
  return FH_OK;
}

}  // extern "C"
