// C-ABI entry points (include/fasthip.h).
//
// Argument validation follows CompressImageData (Core/src/TexComp.cpp:427-525)
// and reports the same conditions as status codes; the message the reference
// prints to stderr is available from fh_last_error().
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fasthip.h"
#include "internal.h"

namespace fh {
namespace {

thread_local std::string t_last_error;

fh_status fail(fh_status s, const char *msg) {
  t_last_error = msg;
  return s;
}

fh_status fail_hip(hipError_t e, const char *what) {
  char buf[256];
  snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
  t_last_error = buf;
  return FH_ERR_HIP;
}

// Per-device state: a growable scratch pool, the process-wide watermark
// counter (the reference's gWMVal, initially 0xFFFFFFFF) and the BC7
// single-colour table.  The counter, the table and the no-allocation work
// area live in the code object (static_mem.hip); `mem` holds their device
// addresses.
//
// Calls on different streams share the pool and the counter.  Each call
// that uses them records `last_use` on its stream after its last launch, and
// the next call's stream waits for it first, so calls on one device run
// their encodes in call order whatever streams they use.
struct DeviceState {
  std::mutex mu;
  int dev = 0;
  StaticDeviceMemory mem;
  void *scratch = nullptr;
  size_t scratch_size = 0;
  uint32_t *counter = nullptr;  // device words: mem.wm
  uint32_t *sctab = nullptr;    // mem.sctab
  hipEvent_t last_use = nullptr;
  bool last_use_recorded = false;
  uint8_t *in_buf = nullptr;
  size_t in_size = 0;
  uint8_t *out_buf = nullptr;
  size_t out_size = 0;
  unsigned long long *sse = nullptr;  // 2 words, fh_psnr's total
  void *stats_buf = nullptr;          // fh_compress_with_stats' records
  size_t stats_size = 0;
  void *sel_buf = nullptr;            // fh_compress_with_selection's selections
  size_t sel_size = 0;
  unsigned long long *work_counts = nullptr;  // fh_bc7_work_counts' two counters
  bool init = false;
};

std::atomic<int> g_work_counting{0};

std::mutex g_devices_mu;
std::vector<DeviceState *> g_devices;

DeviceState *device_state(int dev) {
  std::lock_guard<std::mutex> g(g_devices_mu);
  if ((int)g_devices.size() <= dev) g_devices.resize(dev + 1, nullptr);
  if (!g_devices[dev]) {
    g_devices[dev] = new DeviceState();
    g_devices[dev]->dev = dev;
  }
  return g_devices[dev];
}

// One-time set-up per device (caller holds st->mu, the device current):
// look up the code-object memory and fill the single-colour table.  Nothing
// here allocates device memory.
hipError_t ensure_init(DeviceState *st) {
  if (st->init) return hipSuccess;
  hipError_t e = static_device_memory(st->dev, &st->mem);
  if (e != hipSuccess) return e;
  static uint32_t tab[kSingleColorEntries];
  static std::once_flag built;
  std::call_once(built, [] { build_bc7_single_color_table(tab); });
  e = hipMemcpy(st->mem.sctab, tab, sizeof(tab), hipMemcpyHostToDevice);
  if (e != hipSuccess) return e;
  st->counter = st->mem.wm;
  st->sctab = st->mem.sctab;
  st->init = true;
  return hipSuccess;
}

bool capturing(hipStream_t stream) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

// Orders this call's use of the shared pool and counter after the previous
// call's (see DeviceState).  Not while the stream is being captured: a graph
// is ordered by its caller.
hipError_t order_after_last_use(DeviceState *st, hipStream_t stream) {
  if (!st->last_use_recorded || capturing(stream)) return hipSuccess;
  return hipStreamWaitEvent(stream, st->last_use, 0);
}

hipError_t record_use(DeviceState *st, hipStream_t stream) {
  if (capturing(stream)) return hipSuccess;
  if (!st->last_use) {
    hipError_t e = hipEventCreateWithFlags(&st->last_use, hipEventDisableTiming);
    if (e != hipSuccess) return e;
  }
  hipError_t e = hipEventRecord(st->last_use, stream);
  if (e == hipSuccess) st->last_use_recorded = true;
  return e;
}

bool aligned(const void *p, size_t a) { return (reinterpret_cast<uintptr_t>(p) % a) == 0; }

// Block counts of decoders and error reductions stay within 32-bit lane
// indices (with room for the last workgroup).
bool decode_dims_ok(uint32_t width, uint32_t height) {
  const uint64_t nb = (uint64_t)((width + 3ull) / 4) * ((height + 3ull) / 4);
  return nb <= 0xFFFFFF00ull;
}

hipError_t grow(void **p, size_t *cur, size_t need) {
  if (*cur >= need) return hipSuccess;
  if (*p) {
    hipError_t e = hipFree(*p);
    if (e != hipSuccess) return e;
    *p = nullptr;
    *cur = 0;
  }
  size_t sz = need + need / 4 + 4096;
  hipError_t e = hipMalloc(p, sz);
  if (e != hipSuccess) return e;
  *cur = sz;
  return hipSuccess;
}

uint32_t block_bytes(uint32_t format) {
  switch (format) {
    case FH_FORMAT_DXT1:
    case FH_FORMAT_ETC1:
      return 8;
    case FH_FORMAT_DXT5:
    case FH_FORMAT_BPTC:
      return 16;
    default:
      return 0;
  }
}

// Decoder formats: the four codecs, and BPTC | FH_DECODE_SPEC (the BC7
// specification's decode); -1 for anything else.
int decode_format(uint32_t format) {
  if (format == (FH_FORMAT_BPTC | FH_DECODE_SPEC)) return kFormatBptcSpec;
  return block_bytes(format) ? (int)format : -1;
}

// The block range a (sub-)job covers in raster order, following the loop of
// BPTCC::Compress / DXTC::CompressImageDXT1 (Compressor.cpp:1476-1482):
// rows j = YStart .. min(YEnd, H-4) step 4, the first row from XStart, the
// row j == YEnd only up to XEnd.  Coordinates must be block-aligned.
fh_status job_range(const fh_job *j, uint32_t *first, uint32_t *count) {
  const uint32_t bw = j->width / 4;
  if ((j->x_start | j->y_start | j->x_end) & 3u)
    return fail(FH_ERR_INVALID_ARGUMENT, "job rectangle must be block aligned");
  if (j->y_end < j->height && (j->y_end & 3u))
    return fail(FH_ERR_INVALID_ARGUMENT, "job rectangle must be block aligned");
  if (j->x_start > j->width || j->x_end > j->width || j->y_start > j->height ||
      j->y_end > j->height)
    return fail(FH_ERR_INVALID_ARGUMENT, "job rectangle outside the image");
  const uint32_t start = (j->y_start / 4) * bw + j->x_start / 4;
  const uint32_t end_y = j->y_end < j->height - 4 ? j->y_end : j->height - 4;
  if (j->y_start > end_y) {
    *first = start;
    *count = 0;
    return FH_OK;
  }
  // last processed row
  const uint32_t last_row = j->y_start + ((end_y - j->y_start) / 4) * 4;
  uint32_t end;
  if (last_row == j->y_end) {
    end = (last_row / 4) * bw + (j->x_end + 3) / 4;
    if (last_row == j->y_start && j->x_end <= j->x_start) end = start;
  } else {
    end = (last_row / 4 + 1) * bw;
  }
  *first = start;
  *count = end > start ? end - start : 0;
  return FH_OK;
}

fh_status validate(const fh_job *j, bool device) {
  (void)device;
  if (!j) return fail(FH_ERR_INVALID_ARGUMENT, "null job");
  if (!j->in || (uint64_t)j->width * j->height == 0)
    return fail(FH_ERR_NO_DATA, "No data sent to compress!");
  if (block_bytes(j->format) == 0) return fail(FH_ERR_UNKNOWN_FORMAT, "Unknown compression format");
  if (j->width % 4 || j->height % 4)
    return fail(FH_ERR_BLOCK_DIMS,
                "ERROR - CompressImageData: width or height is not multiple of block dimension");
  const size_t need = (size_t)(j->width / 4) * (j->height / 4) * block_bytes(j->format);
  if (!j->out || j->out_size < need)
    return fail(FH_ERR_BUFFER_TOO_SMALL, "Not enough space for compressed data!");
  if (j->format == FH_FORMAT_BPTC && j->error_metric > 1)
    return fail(FH_ERR_INVALID_ARGUMENT, "unknown error metric");
  if (j->format == FH_FORMAT_BPTC && j->error_metric == 1 && (j->flags & FH_FLAG_LEARNED))
    return fail(FH_ERR_UNSUPPORTED,
                "the learned search ranks for the uniform metric: the nonuniform metric "
                "runs the reference search only");
  // a mask without any of modes 0-7 leaves no candidate for any searched
  // block (the reference asserts, Compressor.cpp:1857)
  if (j->format == FH_FORMAT_BPTC && j->block_modes && !(j->block_modes & 0xFFu))
    return fail(FH_ERR_INVALID_ARGUMENT, "block_modes selects no BC7 mode");
  if ((j->flags & FH_FLAG_CALLER_WATERMARK) && !j->wm_counter)
    return fail(FH_ERR_INVALID_ARGUMENT, "FH_FLAG_CALLER_WATERMARK needs wm_counter");
  if ((j->flags & FH_FLAG_LEARNED) && j->format == FH_FORMAT_BPTC) {
    if (!j->ranker) return fail(FH_ERR_INVALID_ARGUMENT, "FH_FLAG_LEARNED needs a ranker");
    if (j->ranker_topk > 8) return fail(FH_ERR_INVALID_ARGUMENT, "ranker_topk must be 1..8");
    int dev = -1;
    if (hipGetDevice(&dev) == hipSuccess && dev != j->ranker->net.device)
      return fail(FH_ERR_INVALID_ARGUMENT, "ranker belongs to another device");
  }
  return FH_OK;
}

// Bytes of device scratch one BC7 call of up to `blocks` blocks takes: the
// watermark scan's prefixes, then the work area.
size_t bc7_call_bytes(uint32_t blocks, bool learned, bool nonuniform, size_t *work_off) {
  const uint32_t nchunks = (blocks + 255) / 256;
  const size_t prefix_bytes = ((size_t)blocks + 255) & ~(size_t)255;
  *work_off = (prefix_bytes + (size_t)nchunks * 4 + 255) & ~(size_t)255;
  return *work_off + bc7_work_bytes(blocks, learned, nonuniform) + 256;
}

// Runs the encoder on device buffers.  Caller holds st->mu.  With
// fixed_work, BC7 works in that area (fixed_size bytes, large enough for the
// job) and nothing is allocated; otherwise in the device's growable pool.
fh_status run_device(DeviceState *st, const fh_job *j, const uint8_t *din, uint8_t *dout,
                     hipStream_t stream, uint32_t first, uint32_t count,
                     fh_bc7_block_stats *dstats = nullptr,
                     const fh_shape_selection *dsels = nullptr,
                     uint8_t *fixed_work = nullptr, size_t fixed_size = 0) {
  BlockRange r;
  r.img = reinterpret_cast<const uint32_t *>(din);
  r.out = dout;
  r.width = j->width;
  r.blocks_wide = j->width / 4;
  r.first_block = first;
  r.num_blocks = count;
  if (count == 0) return FH_OK;
  hipError_t e = hipSuccess;
  switch (j->format) {
    case FH_FORMAT_BPTC: {
      const uint32_t chunk = kBc7MaxBlocksPerCall;
      const uint32_t per = count < chunk ? count : chunk;
      const size_t prefix_bytes = ((size_t)per + 255) & ~(size_t)255;
      const bool learned = (j->flags & FH_FLAG_LEARNED) != 0;
      const bool nonuniform = j->error_metric == 1 && !dstats;
      size_t work_off = 0;
      const size_t need = bc7_call_bytes(per, learned, nonuniform, &work_off);
      uint8_t *sbase = fixed_work;
      if (fixed_work) {
        if (need > fixed_size) return fail(FH_ERR_INVALID_ARGUMENT, "job too large for the work area");
      } else {
        e = order_after_last_use(st, stream);
        if (e != hipSuccess) return fail_hip(e, "stream ordering");
        if (st->scratch_size < need && st->last_use_recorded) {
          // the pool is about to be replaced: the previous call must be done with it
          e = hipEventSynchronize(st->last_use);
          if (e != hipSuccess) return fail_hip(e, "previous encode");
        }
        e = grow(&st->scratch, &st->scratch_size, need);
        if (e != hipSuccess) return fail_hip(e, "scratch allocation");
        sbase = reinterpret_cast<uint8_t *>(st->scratch);
      }
      Bc7Scratch s;
      s.local_prefix = sbase;
      s.chunk_prefix = reinterpret_cast<uint32_t *>(sbase + prefix_bytes);
      void *work = sbase + work_off;
      s.counter = st->counter;
      uint32_t caller_c = 0;
      if (j->flags & FH_FLAG_CALLER_WATERMARK) {
        // caller's counter lives in host memory: stage it in counter[1]
        s.counter = st->counter + 1;
        caller_c = *j->wm_counter;
        e = hipMemcpyAsync(s.counter, &caller_c, 4, hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) return fail_hip(e, "watermark upload");
      }
      Bc7Params p;
      p.quality = j->quality;
      p.block_modes = j->block_modes ? (j->block_modes & 0xFFu) : 0xFFu;
      p.seed = j->seed;
      p.block_id_base = 0;
      p.q0_guard = (j->flags & FH_FLAG_NO_Q0_GUARD) ? 0u : 1u;
      p.ranker = learned ? &j->ranker->net : nullptr;
      p.topk = j->ranker_topk ? j->ranker_topk : 4;
      p.threshold = j->ranker_threshold;
      p.wg_per_cu = 0;
      p.anneal_unroll = 4;  // annealing steps between refills: 4 measured fastest for both searches
      p.stats = nullptr;
      p.selections = nullptr;
      p.work_counts = nullptr;
      if (g_work_counting.load() && !fixed_work && !nonuniform) {
        if (!st->work_counts) {
          e = hipMalloc(reinterpret_cast<void **>(&st->work_counts), 16);
          if (e != hipSuccess) return fail_hip(e, "work counter allocation");
        }
        e = fill_u32(reinterpret_cast<uint32_t *>(st->work_counts), 4, 0u, stream);
        if (e != hipSuccess) return fail_hip(e, "work counter reset");
        p.work_counts = st->work_counts;
      }
      // kErrorMetrics (Compressor.cpp:205-208)
      p.nonuniform = nonuniform ? 1u : 0u;
      p.metric[0] = nonuniform ? sqrtf(0.3f) : 1.0f;
      p.metric[1] = nonuniform ? sqrtf(0.56f) : 1.0f;
      p.metric[2] = nonuniform ? sqrtf(0.11f) : 1.0f;
      p.metric[3] = 1.0f;
      // jobs larger than one call's work area run as consecutive calls on the
      // stream; the watermark counter carries across them in order
      for (uint32_t done = 0; done < count; done += chunk) {
        BlockRange rc = r;
        rc.first_block = first + done;
        rc.num_blocks = count - done < chunk ? count - done : chunk;
        s.num_chunks = (rc.num_blocks + 255) / 256;
        if (dstats) p.stats = dstats + done;
        if (dsels) p.selections = dsels + done;
        e = bc7_encode(rc, p, s, work, stream);
        if (e != hipSuccess) return fail_hip(e, "BC7 encode launch");
      }
      if (!fixed_work) {
        e = record_use(st, stream);
        if (e != hipSuccess) return fail_hip(e, "stream ordering");
      }
      if (j->flags & FH_FLAG_CALLER_WATERMARK) {
        e = hipMemcpyAsync(&caller_c, s.counter, 4, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return fail_hip(e, "watermark download");
        *j->wm_counter = caller_c;
      }
      return FH_OK;
    }
    case FH_FORMAT_DXT1:
    case FH_FORMAT_DXT5:
      e = dxt_encode(r, j->format == FH_FORMAT_DXT5, stream);
      if (e != hipSuccess) return fail_hip(e, "DXT encode launch");
      return FH_OK;
    case FH_FORMAT_ETC1:
      e = etc1_encode(r, stream);
      if (e != hipSuccess) return fail_hip(e, "ETC1 encode launch");
      return FH_OK;
  }
  return fail(FH_ERR_UNKNOWN_FORMAT, "Unknown compression format");
}

int current_device(fh_status *s) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    *s = fail(FH_ERR_NO_DEVICE, "no HIP device available");
    return -1;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    *s = fail(FH_ERR_NO_DEVICE, "hipGetDevice failed");
    return -1;
  }
  *s = FH_OK;
  return dev;
}

// Candidates of one caller selection, counted as the select kernel builds
// them (CompressClusters' loop, Compressor.cpp:1771-1809): rotation modes 8
// and 4 slots, a single-subset mode once per entry only when annealing.
int selection_candidates(const fh_shape_selection &g, uint32_t block_modes, uint32_t quality) {
  static const int kSubsets[8] = {3, 2, 3, 2, 1, 1, 1, 2};
  static const int kSlots[8] = {1, 1, 1, 1, 8, 4, 1, 1};
  uint32_t modes = g.selected_modes & block_modes;
  uint32_t nsh = g.num_shapes < 5 ? g.num_shapes : 5;
  if (nsh == 0) {
    nsh = 1;
    modes &= ~0x8Fu;  // modes 0, 1, 2, 3 and 7
  }
  int n = 0;
  for (int mode = 0; mode < 8; mode++) {
    if (!((modes >> mode) & 1)) continue;
    for (uint32_t e = 0; e < nsh; e++) {
      const int nsub = kSubsets[mode];
      if (nsub != 1 && (uint32_t)nsub != g.shapes[e].num_partitions) continue;
      if (mode == 0 && g.shapes[e].index >= 16) continue;
      if (nsub == 1 && e > 0 && quality == 0) continue;
      n += kSlots[mode];
    }
  }
  return n;
}

// Whether block `b` (raster index) of a host image takes the selection
// function's path: neither one colour nor fully transparent
// (CompressBC7Block, Compressor.cpp:1819-1860).
bool searched_block(const fh_job *j, uint32_t b) {
  const uint32_t bw = j->width / 4;
  const uint32_t *img = reinterpret_cast<const uint32_t *>(j->in);
  const uint32_t x = (b % bw) * 4, y = (b / bw) * 4;
  const uint32_t p0 = img[(size_t)y * j->width + x];
  bool same = true, transparent = true;
  for (uint32_t r = 0; r < 4; r++)
    for (uint32_t c = 0; c < 4; c++) {
      const uint32_t p = img[(size_t)(y + r) * j->width + x + c];
      same = same && p == p0;
      transparent = transparent && (p >> 24) == 0;
    }
  return !same && !transparent;
}

// Caller selections: shape indices inside the tables and at most 31
// candidates per block; with a host image (first = the job's first block),
// every searched block must get at least one candidate, where the reference
// asserts (Compressor.cpp:1857).
fh_status check_selections(const fh_shape_selection *sels, uint32_t count, const fh_job *j,
                           bool host_image = false, uint32_t first = 0) {
  const uint32_t modes = j->block_modes ? (j->block_modes & 0xFFu) : 0xFFu;
  for (uint32_t k = 0; k < count; k++) {
    const fh_shape_selection &g = sels[k];
    const uint32_t nsh = g.num_shapes < 5 ? g.num_shapes : 5;
    for (uint32_t e = 0; e < nsh; e++) {
      const uint32_t parts = g.shapes[e].num_partitions;
      if ((parts == 2 || parts == 3) && g.shapes[e].index >= 64)
        return fail(FH_ERR_INVALID_ARGUMENT, "shape index beyond the partition tables");
    }
    const int n = selection_candidates(g, modes, j->quality);
    if (n > 31) return fail(FH_ERR_UNSUPPORTED, "a shape selection with more than 31 candidates");
    if (n == 0 && host_image && searched_block(j, first + k))
      return fail(FH_ERR_INVALID_ARGUMENT,
                  "a shape selection that leaves a searched block no mode to try");
  }
  return FH_OK;
}

// The statistics and shape-selection entry points take BPTC jobs of the
// reference search only.
fh_status stats_job_ok(const fh_job *j) {
  if (j && j->format != FH_FORMAT_BPTC && block_bytes(j->format) != 0)
    return fail(FH_ERR_UNSUPPORTED, "statistics and shape selections are BPTC only");
  if (j && (j->flags & FH_FLAG_LEARNED))
    return fail(FH_ERR_INVALID_ARGUMENT, "not combined with the learned search");
  return FH_OK;
}

// fh_compress and fh_compress_with_stats on host buffers: stage the image,
// encode, copy the blocks (and statistics) back.  Caller-facing checks as in
// CompressImageData.
fh_status compress_host(const fh_job *j, bool with_stats, fh_bc7_block_stats *stats,
                        size_t stats_count, const fh_shape_selection *sels = nullptr,
                        size_t sels_count = 0) {
  fh_status s = validate(j, false);
  if (s != FH_OK) return s;
  uint32_t first = 0, count = 0;
  s = job_range(j, &first, &count);
  if (s != FH_OK) return s;
  if (with_stats && count && (!stats || stats_count < count))
    return fail(FH_ERR_BUFFER_TOO_SMALL, "statistics buffer too small");
  if (sels) {
    if (sels_count < count) return fail(FH_ERR_BUFFER_TOO_SMALL, "too few shape selections");
    s = check_selections(sels, count, j, true, first);
    if (s != FH_OK) return s;
  }
  const int dev = current_device(&s);
  if (dev < 0) return s;
  DeviceState *st = device_state(dev);
  std::lock_guard<std::mutex> g(st->mu);
  hipError_t e = ensure_init(st);
  if (e != hipSuccess) return fail_hip(e, "device initialisation");
  const size_t in_bytes = (size_t)j->width * j->height * 4;
  const uint32_t bb = block_bytes(j->format);
  const size_t out_bytes = (size_t)(j->width / 4) * (j->height / 4) * bb;
  e = grow(reinterpret_cast<void **>(&st->in_buf), &st->in_size, in_bytes);
  if (e == hipSuccess)
    e = grow(reinterpret_cast<void **>(&st->out_buf), &st->out_size, out_bytes);
  const size_t stats_bytes = with_stats ? (size_t)count * sizeof(fh_bc7_block_stats) : 0;
  if (e == hipSuccess && stats_bytes) e = grow(&st->stats_buf, &st->stats_size, stats_bytes);
  const size_t sel_bytes = sels ? (size_t)count * sizeof(fh_shape_selection) : 0;
  if (e == hipSuccess && sel_bytes) e = grow(&st->sel_buf, &st->sel_size, sel_bytes);
  if (e != hipSuccess) return fail_hip(e, "device buffer allocation");
  fh_bc7_block_stats *dstats =
      stats_bytes ? reinterpret_cast<fh_bc7_block_stats *>(st->stats_buf) : nullptr;
  hipStream_t stream = nullptr;
  e = hipMemcpyAsync(st->in_buf, j->in, in_bytes, hipMemcpyHostToDevice, stream);
  if (e == hipSuccess && sel_bytes)
    e = hipMemcpyAsync(st->sel_buf, sels, sel_bytes, hipMemcpyHostToDevice, stream);
  if (e != hipSuccess) return fail_hip(e, "input upload");
  s = run_device(st, j, st->in_buf, st->out_buf, stream, first, count, dstats,
                 sel_bytes ? reinterpret_cast<const fh_shape_selection *>(st->sel_buf) : nullptr);
  if (s != FH_OK) return s;
  if (dstats) {
    e = hipMemcpyAsync(stats, dstats, stats_bytes, hipMemcpyDeviceToHost, stream);
    if (e != hipSuccess) return fail_hip(e, "statistics download");
  }
  if (count) {
    e = hipMemcpyAsync(j->out + (size_t)first * bb, st->out_buf + (size_t)first * bb,
                       (size_t)count * bb, hipMemcpyDeviceToHost, stream);
    if (e != hipSuccess) return fail_hip(e, "output download");
  }
  e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return fail_hip(e, "encode");
  return FH_OK;
}

}  // namespace

void set_error(const char *msg) { t_last_error = msg; }

uint32_t *device_wm_counter(int dev) {
  DeviceState *st = device_state(dev);
  std::lock_guard<std::mutex> g(st->mu);
  if (ensure_init(st) != hipSuccess) return nullptr;
  return st->counter;
}

const uint32_t *bc7_single_color_table(int device) {
  DeviceState *st = device_state(device);
  return st->sctab;
}

}  // namespace fh

using namespace fh;

extern "C" {

fh_status fh_compress(const fh_job *j) { return compress_host(j, false, nullptr, 0); }

// CompressWithStats under the nonuniform metric is not built: its shape scan
// logs the uniform estimates only.
fh_status stats_metric_ok(const fh_job *j) {
  if (j && j->error_metric == 1)
    return fail(FH_ERR_UNSUPPORTED,
                "CompressWithStats runs the uniform metric only (the nonuniform one runs "
                "through fh_compress / fh_compress_device)");
  return FH_OK;
}

fh_status fh_compress_with_stats(const fh_job *j, fh_bc7_block_stats *stats, size_t stats_count) {
  fh_status s = stats_job_ok(j);
  if (s != FH_OK) return s;
  s = stats_metric_ok(j);
  if (s != FH_OK) return s;
  return compress_host(j, true, stats, stats_count);
}

fh_status fh_compress_with_stats_device(const fh_job *j, fh_bc7_block_stats *stats,
                                        size_t stats_count, void *hip_stream) {
  fh_status s = stats_job_ok(j);
  if (s != FH_OK) return s;
  s = stats_metric_ok(j);
  if (s != FH_OK) return s;
  s = validate(j, true);
  if (s != FH_OK) return s;
  uint32_t first = 0, count = 0;
  s = job_range(j, &first, &count);
  if (s != FH_OK) return s;
  if (count && (!stats || stats_count < count))
    return fail(FH_ERR_BUFFER_TOO_SMALL, "statistics buffer too small");
  if (!aligned(stats, alignof(fh_bc7_block_stats)))
    return fail(FH_ERR_INVALID_ARGUMENT, "statistics records must be 8-byte aligned");
  const int dev = current_device(&s);
  if (dev < 0) return s;
  DeviceState *st = device_state(dev);
  std::lock_guard<std::mutex> g(st->mu);
  hipError_t e = ensure_init(st);
  if (e != hipSuccess) return fail_hip(e, "device initialisation");
  return run_device(st, j, j->in, j->out, reinterpret_cast<hipStream_t>(hip_stream), first,
                    count, stats);
}

fh_status fh_compress_with_selection(const fh_job *j, const fh_shape_selection *selections,
                                     size_t selections_count) {
  const fh_status s = stats_job_ok(j);
  if (s != FH_OK) return s;
  if (!selections) return fail(FH_ERR_INVALID_ARGUMENT, "null shape selections");
  return compress_host(j, false, nullptr, 0, selections, selections_count);
}

fh_status fh_compress_with_selection_device(const fh_job *j, const fh_shape_selection *selections,
                                            size_t selections_count, void *hip_stream) {
  fh_status s = stats_job_ok(j);
  if (s != FH_OK) return s;
  if (!selections) return fail(FH_ERR_INVALID_ARGUMENT, "null shape selections");
  if (!aligned(selections, alignof(fh_shape_selection)))
    return fail(FH_ERR_INVALID_ARGUMENT, "shape selections must be 4-byte aligned");
  s = validate(j, true);
  if (s != FH_OK) return s;
  uint32_t first = 0, count = 0;
  s = job_range(j, &first, &count);
  if (s != FH_OK) return s;
  if (selections_count < count) return fail(FH_ERR_BUFFER_TOO_SMALL, "too few shape selections");
  const int dev = current_device(&s);
  if (dev < 0) return s;
  DeviceState *st = device_state(dev);
  std::lock_guard<std::mutex> g(st->mu);
  hipError_t e = ensure_init(st);
  if (e != hipSuccess) return fail_hip(e, "device initialisation");
  return run_device(st, j, j->in, j->out, reinterpret_cast<hipStream_t>(hip_stream), first,
                    count, nullptr, selections);
}

fh_status fh_bc7_stats_format(const fh_bc7_block_stats *st, uint64_t block_index, char *out,
                              size_t out_capacity, size_t *out_size) {
  if (!st || !out_size) return fail(FH_ERR_INVALID_ARGUMENT, "null statistics or size");
  if (st->num_two_shape > 64 || st->num_three_shape > 64)
    return fail(FH_ERR_INVALID_ARGUMENT, "more than 64 shape estimates");
  // kBlockStatString (Compressor.cpp:106-131) spells the mode numbers out
  static const char *const kNum[8] = {"Zero", "One", "Two", "Three",
                                      "Four", "Five", "Six", "Seven"};
  std::string text;
  char line[192];
  // BlockLogger prefixes "<index>: " to each "<name> -- <value>\n" that
  // PrintStat builds with an ostream (doubles: default format, i.e. %g)
  auto add_d = [&](const char *name, const char *suffix, double v) {
    snprintf(line, sizeof(line), "%llu: BlockStat_%s%s -- %g\n", (unsigned long long)block_index,
             name, suffix, v);
    text += line;
  };
  auto add_i = [&](const char *name, int v) {
    snprintf(line, sizeof(line), "%llu: BlockStat_%s -- %d\n", (unsigned long long)block_index,
             name, v);
    text += line;
  };
  for (uint32_t k = 0; k < st->num_two_shape; k++) add_d("TwoShapeEstimate", "", st->two_shape[k]);
  for (uint32_t k = 0; k < st->num_three_shape; k++)
    add_d("ThreeShapeEstimate", "", st->three_shape[k]);
  add_i("Path", st->path);
  add_i("Mode", st->mode);
  char name[32];
  for (int m = 0; m < 8; m++) {
    snprintf(name, sizeof(name), "Mode%s", kNum[m]);
    add_d(name, "Estimate", st->mode_estimate[m]);
    add_d(name, "Error", st->mode_error[m]);
  }
  *out_size = text.size();
  if (!out) return FH_OK;
  if (out_capacity <= text.size()) return fail(FH_ERR_BUFFER_TOO_SMALL, "text buffer too small");
  memcpy(out, text.c_str(), text.size() + 1);
  return FH_OK;
}

fh_status fh_compress_device(const fh_job *j, void *hip_stream) {
  fh_status s = validate(j, true);
  if (s != FH_OK) return s;
  uint32_t first = 0, count = 0;
  s = job_range(j, &first, &count);
  if (s != FH_OK) return s;
  const int dev = current_device(&s);
  if (dev < 0) return s;
  DeviceState *st = device_state(dev);
  std::lock_guard<std::mutex> g(st->mu);
  hipError_t e = ensure_init(st);
  if (e != hipSuccess) return fail_hip(e, "device initialisation");
  return run_device(st, j, j->in, j->out, reinterpret_cast<hipStream_t>(hip_stream), first,
                    count);
}

fh_status fh_compress_device_noalloc(const fh_job *j, void *hip_stream) {
  fh_status s = validate(j, true);
  if (s != FH_OK) return s;
  if (j->format != FH_FORMAT_BPTC)
    return fail(FH_ERR_UNSUPPORTED,
                "fh_compress_device_noalloc encodes BPTC only (DXT/ETC1 use no work area: "
                "call fh_compress_device)");
  if (j->flags & (FH_FLAG_LEARNED | FH_FLAG_CALLER_WATERMARK))
    return fail(FH_ERR_UNSUPPORTED,
                "fh_compress_device_noalloc takes neither FH_FLAG_LEARNED nor "
                "FH_FLAG_CALLER_WATERMARK");
  uint32_t first = 0, count = 0;
  s = job_range(j, &first, &count);
  if (s != FH_OK) return s;
  if (count > kNoAllocMaxBlocks)
    return fail(FH_ERR_INVALID_ARGUMENT,
                "job larger than the no-allocation work area (at most 16384 blocks, "
                "e.g. 512x512 texels; split it into sub-jobs)");
  const int dev = current_device(&s);
  if (dev < 0) return s;
  DeviceState *st = device_state(dev);
  std::lock_guard<std::mutex> g(st->mu);
  hipError_t e = ensure_init(st);
  if (e != hipSuccess) return fail_hip(e, "device initialisation");
  size_t work_off = 0;
  if (bc7_call_bytes(kNoAllocMaxBlocks, false, j->error_metric == 1, &work_off) >
      st->mem.scratch_bytes)
    return fail(FH_ERR_HIP, "no-allocation work area smaller than its block limit needs");
  return run_device(st, j, j->in, j->out, reinterpret_cast<hipStream_t>(hip_stream), first,
                    count, nullptr, nullptr, st->mem.scratch, st->mem.scratch_bytes);
}

fh_status fh_decompress(uint32_t format, const uint8_t *in, size_t in_size, uint8_t *out,
                        uint32_t width, uint32_t height) {
  if (!in || !out || width == 0 || height == 0)
    return fail(FH_ERR_INVALID_ARGUMENT, "null buffer or empty image");
  const int df = decode_format(format);
  if (df < 0) return fail(FH_ERR_UNKNOWN_FORMAT, "Unknown compression format");
  const uint32_t bb = block_bytes(format & 0xFFu);
  if (!decode_dims_ok(width, height)) return fail(FH_ERR_INVALID_ARGUMENT, "image too large");
  const size_t need = (size_t)((width + 3) / 4) * ((height + 3) / 4) * bb;
  if (in_size < need) return fail(FH_ERR_BUFFER_TOO_SMALL, "compressed buffer too small");
  fh_status s;
  const int dev = current_device(&s);
  if (dev < 0) return s;
  DeviceState *st = device_state(dev);
  std::lock_guard<std::mutex> g(st->mu);
  hipError_t e = ensure_init(st);
  if (e != hipSuccess) return fail_hip(e, "device initialisation");
  const size_t out_bytes = (size_t)width * height * 4;
  e = grow(reinterpret_cast<void **>(&st->in_buf), &st->in_size, need);
  if (e == hipSuccess)
    e = grow(reinterpret_cast<void **>(&st->out_buf), &st->out_size, out_bytes);
  if (e != hipSuccess) return fail_hip(e, "device buffer allocation");
  e = hipMemcpy(st->in_buf, in, need, hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail_hip(e, "input upload");
  e = texture_decode((uint32_t)df, st->in_buf, reinterpret_cast<uint32_t *>(st->out_buf), width,
                     height, nullptr);
  if (e != hipSuccess) return fail_hip(e, "decode launch");
  e = hipMemcpy(out, st->out_buf, out_bytes, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return fail_hip(e, "output download");
  return FH_OK;
}

fh_status fh_decompress_device(uint32_t format, const uint8_t *in, uint8_t *out, uint32_t width,
                               uint32_t height, void *hip_stream) {
  if (!in || !out) return fail(FH_ERR_INVALID_ARGUMENT, "null buffer");
  const int df = decode_format(format);
  if (df < 0) return fail(FH_ERR_UNKNOWN_FORMAT, "Unknown compression format");
  if (!decode_dims_ok(width, height)) return fail(FH_ERR_INVALID_ARGUMENT, "image too large");
  if (!aligned(in, block_bytes(format & 0xFFu)) || !aligned(out, 4))
    return fail(FH_ERR_INVALID_ARGUMENT, "misaligned buffer (blocks: block size, pixels: 4 B)");
  hipError_t e = texture_decode((uint32_t)df, in, reinterpret_cast<uint32_t *>(out), width, height,
                                reinterpret_cast<hipStream_t>(hip_stream));
  if (e != hipSuccess) return fail_hip(e, "decode launch");
  return FH_OK;
}

fh_status fh_block_error_device(uint32_t format, const uint32_t *original, const uint8_t *encoded,
                                uint32_t width, uint32_t height,
                                unsigned long long *block_errors, unsigned long long *sse,
                                void *hip_stream) {
  if (!original || !encoded || !sse) return fail(FH_ERR_INVALID_ARGUMENT, "null buffer");
  const int df = decode_format(format);
  if (df < 0) return fail(FH_ERR_UNKNOWN_FORMAT, "Unknown compression format");
  if (!decode_dims_ok(width, height)) return fail(FH_ERR_INVALID_ARGUMENT, "image too large");
  if (!aligned(encoded, block_bytes(format & 0xFFu)) || !aligned(original, 4) ||
      !aligned(sse, 8) || !aligned(block_errors, 8))
    return fail(FH_ERR_INVALID_ARGUMENT, "misaligned buffer");
  hipError_t e = block_error(df, original, encoded, width, height, block_errors, sse,
                             reinterpret_cast<hipStream_t>(hip_stream));
  if (e != hipSuccess) return fail_hip(e, "error launch");
  return FH_OK;
}

fh_status fh_image_error_device(const uint32_t *a, const uint32_t *b, uint32_t width,
                                uint32_t height, unsigned long long *block_errors,
                                unsigned long long *sse, void *hip_stream) {
  if (!a || !b || !sse) return fail(FH_ERR_INVALID_ARGUMENT, "null buffer");
  if (!decode_dims_ok(width, height)) return fail(FH_ERR_INVALID_ARGUMENT, "image too large");
  if (!aligned(a, 4) || !aligned(b, 4) || !aligned(sse, 8) || !aligned(block_errors, 8))
    return fail(FH_ERR_INVALID_ARGUMENT, "misaligned buffer");
  hipError_t e = block_error(kFormatRgba8, a, b, width, height, block_errors, sse,
                             reinterpret_cast<hipStream_t>(hip_stream));
  if (e != hipSuccess) return fail_hip(e, "error launch");
  return FH_OK;
}

fh_status fh_mip_level_device(const uint32_t *src, uint32_t width, uint32_t height, uint32_t *dst,
                              void *hip_stream) {
  if (!src || !dst || width == 0 || height == 0)
    return fail(FH_ERR_INVALID_ARGUMENT, "null buffer or empty image");
  if (width == 1 && height == 1) return fail(FH_ERR_INVALID_ARGUMENT, "a 1x1 level has no next level");
  if (!aligned(src, 4) || !aligned(dst, 4)) return fail(FH_ERR_INVALID_ARGUMENT, "misaligned buffer");
  hipError_t e = mip_level(src, width, height, dst, reinterpret_cast<hipStream_t>(hip_stream));
  if (e != hipSuccess) return fail_hip(e, "mip launch");
  return FH_OK;
}

fh_status fh_pad4_device(const uint32_t *src, uint32_t width, uint32_t height, uint32_t *dst,
                         void *hip_stream) {
  if (!src || !dst || width == 0 || height == 0)
    return fail(FH_ERR_INVALID_ARGUMENT, "null buffer or empty image");
  if (width > 0xFFFFFFFCu || height > 0xFFFFFFFCu)
    return fail(FH_ERR_INVALID_ARGUMENT, "image too large");
  if (!aligned(src, 4) || !aligned(dst, 4)) return fail(FH_ERR_INVALID_ARGUMENT, "misaligned buffer");
  hipError_t e = pad4(src, width, height, dst, reinterpret_cast<hipStream_t>(hip_stream));
  if (e != hipSuccess) return fail_hip(e, "pad launch");
  return FH_OK;
}

double fh_psnr_from_sse(const unsigned long long sse[2], uint32_t width, uint32_t height) {
  // sse is 255^2 times the reference's sum of squared differences
  const double total = (double)sse[1] * 18446744073709551616.0 + (double)sse[0];
  const double mse = total / (255.0 * 255.0) / ((double)width * (double)height);
  return 10.0 * log10(3.0 * 255.0 * 255.0 / mse);
}

fh_status fh_psnr(const uint8_t *a, const uint8_t *b, uint32_t width, uint32_t height,
                  double *psnr) {
  if (!a || !b || !psnr || width == 0 || height == 0)
    return fail(FH_ERR_INVALID_ARGUMENT, "null buffer or empty image");
  if (!decode_dims_ok(width, height)) return fail(FH_ERR_INVALID_ARGUMENT, "image too large");
  fh_status s;
  const int dev = current_device(&s);
  if (dev < 0) return s;
  DeviceState *st = device_state(dev);
  std::lock_guard<std::mutex> g(st->mu);
  hipError_t e = ensure_init(st);
  if (e != hipSuccess) return fail_hip(e, "device initialisation");
  const size_t bytes = (size_t)width * height * 4;
  e = grow(reinterpret_cast<void **>(&st->in_buf), &st->in_size, bytes);
  if (e == hipSuccess) e = grow(reinterpret_cast<void **>(&st->out_buf), &st->out_size, bytes);
  if (e == hipSuccess && !st->sse) e = hipMalloc(&st->sse, 2 * sizeof(unsigned long long));
  if (e != hipSuccess) return fail_hip(e, "device buffer allocation");
  e = hipMemcpy(st->in_buf, a, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(st->out_buf, b, bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail_hip(e, "image upload");
  e = block_error(kFormatRgba8, reinterpret_cast<const uint32_t *>(st->in_buf), st->out_buf,
                  width, height, nullptr, st->sse, nullptr);
  if (e != hipSuccess) return fail_hip(e, "error launch");
  unsigned long long total[2];
  e = hipMemcpy(total, st->sse, sizeof(total), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return fail_hip(e, "error download");
  *psnr = fh_psnr_from_sse(total, width, height);
  return FH_OK;
}

fh_status fh_compressed_size(uint32_t format, uint32_t width, uint32_t height, size_t *size) {
  if (!size) return fail(FH_ERR_INVALID_ARGUMENT, "null size");
  const uint32_t bb = block_bytes(format);
  if (!bb) return fail(FH_ERR_UNKNOWN_FORMAT, "Unknown compression format");
  *size = (size_t)((width + 3) / 4) * ((height + 3) / 4) * bb;
  return FH_OK;
}

fh_status fh_bc7_exhaustive_labels(const uint32_t *blocks, uint32_t n, int32_t *labels,
                                   unsigned long long *errors, void *hip_stream) {
  if (!blocks || !labels) return fail(FH_ERR_INVALID_ARGUMENT, "null buffer");
  if (n == 0) return FH_OK;
  fh_status s;
  const int dev = current_device(&s);
  if (dev < 0) return s;
  DeviceState *st = device_state(dev);
  std::lock_guard<std::mutex> g(st->mu);
  hipError_t e = ensure_init(st);
  if (e != hipSuccess) return fail_hip(e, "device initialisation");
  const hipStream_t stream = reinterpret_cast<hipStream_t>(hip_stream);
  e = order_after_last_use(st, stream);
  if (e == hipSuccess && st->scratch_size < (size_t)n * 8 + 256 && st->last_use_recorded)
    e = hipEventSynchronize(st->last_use);
  if (e != hipSuccess) return fail_hip(e, "previous encode");
  e = grow(&st->scratch, &st->scratch_size, (size_t)n * 8 + 256);
  if (e != hipSuccess) return fail_hip(e, "scratch allocation");
  // n blocks of 16 pixels, seen as a 4 x 4n image one block wide
  BlockRange r;
  r.img = blocks;
  r.out = nullptr;
  r.width = 4;
  r.blocks_wide = 1;
  r.first_block = 0;
  r.num_blocks = n;
  e = bc7_exhaustive_labels(r, reinterpret_cast<unsigned long long *>(st->scratch), labels,
                            errors, stream);
  if (e != hipSuccess) return fail_hip(e, "label search launch");
  e = record_use(st, stream);
  if (e != hipSuccess) return fail_hip(e, "stream ordering");
  return FH_OK;
}

fh_status fh_bc7_candidate_errors(const uint32_t *blocks, uint32_t n, float *errors,
                                  void *hip_stream) {
  if (!blocks || !errors) return fail(FH_ERR_INVALID_ARGUMENT, "null buffer");
  if (n == 0) return FH_OK;
  if ((uint64_t)n * 512 >= (1ull << 32))
    return fail(FH_ERR_INVALID_ARGUMENT, "at most 2^23 - 1 blocks per call");
  fh_status s;
  const int dev = current_device(&s);
  if (dev < 0) return s;
  DeviceState *st = device_state(dev);
  std::lock_guard<std::mutex> g(st->mu);
  hipError_t e = ensure_init(st);
  if (e != hipSuccess) return fail_hip(e, "device initialisation");
  BlockRange r;
  r.img = blocks;
  r.out = nullptr;
  r.width = 4;
  r.blocks_wide = 1;
  r.first_block = 0;
  r.num_blocks = n;
  e = bc7_candidate_errors(r, errors, reinterpret_cast<hipStream_t>(hip_stream));
  if (e != hipSuccess) return fail_hip(e, "candidate error launch");
  return FH_OK;
}

const char *fh_last_error(void) { return t_last_error.c_str(); }

fh_status fh_set_work_counting(int enable) {
  g_work_counting.store(enable ? 1 : 0);
  return FH_OK;
}

fh_status fh_bc7_work_counts(uint64_t counts[2]) {
  if (!counts) return fail(FH_ERR_INVALID_ARGUMENT, "null counts");
  fh_status s;
  const int dev = current_device(&s);
  if (dev < 0) return s;
  DeviceState *st = device_state(dev);
  std::lock_guard<std::mutex> g(st->mu);
  counts[0] = counts[1] = 0;
  if (!st->work_counts) return FH_OK;
  hipError_t e = hipDeviceSynchronize();
  unsigned long long c[2] = {0, 0};
  if (e == hipSuccess) e = hipMemcpy(c, st->work_counts, sizeof(c), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return fail_hip(e, "work counter read");
  counts[0] = c[0];
  counts[1] = c[1];
  return FH_OK;
}

int fh_device_available(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n > 0;
}

}  // extern "C"
