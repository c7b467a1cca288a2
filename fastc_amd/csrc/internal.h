// Internal interfaces between the C-ABI layer (capi.cpp) and the HIP kernel
// translation units.  Not installed; callers use include/fasthip.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

struct fh_bc7_block_stats;  // include/fasthip.h
struct fh_shape_selection;

namespace fh {

// Launch parameters shared by the block encoders.  Blocks are addressed in
// image raster order; a job covers blocks [first_block, first_block+num_blocks).
struct BlockRange {
  const uint32_t *img;   // device, width*height packed RGBA8
  uint8_t *out;          // device, base of the whole compressed image
  uint32_t width;        // pixels
  uint32_t blocks_wide;  // width / 4
  uint32_t first_block;
  uint32_t num_blocks;
};

struct Ranker;

struct Bc7Params {
  uint32_t quality;      // annealing steps
  uint32_t block_modes;  // mask of allowed BC7 modes
  uint64_t seed;         // annealing stream key
  uint64_t block_id_base;  // added to the raster block index to key the stream
  uint32_t q0_guard;     // keep the quality-0 block where annealing decodes worse
  // learned candidate pruning (null = the reference's BoxSelection search)
  const Ranker *ranker;
  uint32_t topk;         // candidates searched per confident block (1..8)
  float threshold;       // below this top softmax probability: BoxSelection
  uint32_t wg_per_cu;    // persistent workgroups per CU (0 = 16)
  uint32_t anneal_unroll;  // annealing steps between refills (0 = 1)
  // CompressWithStats' block path: one record per block of the range (null =
  // Compress's BoxSelection search); not combined with the ranker
  fh_bc7_block_stats *stats;
  // a caller's shape selection per block of the range (device; null = the
  // library's own search); not combined with the ranker or statistics
  const fh_shape_selection *selections;
  // m_ErrorMetric (BPTCCompressor.h:73-77): 0 the uniform metric, 1 the
  // nonuniform one with per-channel weights `metric` (kErrorMetrics,
  // Compressor.cpp:205-208); not combined with the ranker or statistics
  uint32_t nonuniform;
  float metric[4];
  // work counting (device words, null = off): the annealing's reference
  // QuantizedError point evaluations and this encoder's pixel scores
  unsigned long long *work_counts;
};

// Device scratch needed by the BC7 watermark prefix scan.
struct Bc7Scratch {
  uint8_t *local_prefix;   // per block: solid blocks before it inside its 256-block chunk
  uint32_t *chunk_prefix;  // per chunk: exclusive solid-block count
  uint32_t *counter;       // watermark counter in/out (device word)
  uint32_t num_chunks;
};

// Single-colour endpoint tables (uploaded once per device).
const uint32_t *bc7_single_color_table(int device);

// Device work area bc7_encode needs for num_blocks blocks (candidate
// records, queues and annealing states, and with the ranker its
// activations).
size_t bc7_work_bytes(uint32_t num_blocks, bool learned, bool nonuniform = false);

// Largest block count one bc7_encode call takes; callers split bigger jobs.
constexpr uint32_t kBc7MaxBlocksPerCall = 1u << 20;

// Entries of the BC7 single-colour endpoint table (tables.cpp).
constexpr int kSingleColorEntries = 5 * 5 * 4 * 256;

// Devices the per-device tables cover.
constexpr int kMaxDevices = 64;

// fh_compress_device_noalloc: jobs of at most this many blocks (a 512x512
// image) run in a fixed per-device work area of kNoAllocScratchBytes.
constexpr uint32_t kNoAllocMaxBlocks = 1u << 14;
constexpr size_t kNoAllocScratchBytes = (size_t)72 << 20;

// Device addresses of the memory kept in the code object (static_mem.hip):
// the watermark counter words, the BC7 single-colour table and the no-alloc
// work area.  Looked up once per device with that device current.
struct StaticDeviceMemory {
  uint32_t *wm = nullptr;
  uint32_t *sctab = nullptr;
  uint8_t *scratch = nullptr;
  size_t scratch_bytes = 0;
};
hipError_t static_device_memory(int dev, StaticDeviceMemory *out);

hipError_t bc7_encode(const BlockRange &r, const Bc7Params &p, const Bc7Scratch &s, void *work,
                      hipStream_t stream);

// Exhaustive quality-0 search over all 285 (mode, slot) candidates of the
// candidate ranker: labels[i] = mode * 64 + slot of the smallest decoded
// premultiplied error (errors[i], optional; Image::ComputePSNR's per-pixel
// term times 255^2).  best: n words of scratch.
hipError_t bc7_exhaustive_labels(const BlockRange &r, unsigned long long *best, int32_t *labels,
                                 unsigned long long *errors, hipStream_t stream);
// Every candidate's decoded error (errs: n x 512 floats in logit order,
// +inf where a logit is no candidate).
hipError_t bc7_candidate_errors(const BlockRange &r, float *errs, hipStream_t stream);

hipError_t dxt_encode(const BlockRange &r, int dxt5, hipStream_t stream);
hipError_t etc1_encode(const BlockRange &r, hipStream_t stream);

// ---- Batches of surfaces (texture levels) --------------------------------
// One surface of a batch, on the device.  Blocks of a format group are
// numbered globally in submission order; surface s owns blocks
// [first, first + nblocks) and block_surf maps a global block to its surface.
struct SurfDesc {
  const uint32_t *img;   // width * height RGBA8
  uint8_t *out;          // the surface's compressed blocks (raster order)
  uint32_t width;        // pixels, a multiple of 4
  uint32_t blocks_wide;
  uint32_t first;        // first global block of the surface
  uint32_t nblocks;
};

// Launch shapes the batch autotuner chooses.
struct LaunchTune {
  uint32_t wg_per_cu = 16;     // persistent workgroups per CU (BC7 set-up / annealing)
  uint32_t anneal_unroll = 1;  // annealing steps between refills (1, 2 or 4)
  uint32_t simple_wg = 256;    // workgroup size of the DXT / ETC1 batch kernels
};

// Sets n words to v with a kernel (safe inside captured graphs).
hipError_t fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t stream);

// The library's process-wide BC7 watermark counter on device dev (capi.cpp),
// initialising the device state if needed; null on failure.
uint32_t *device_wm_counter(int dev);

// Table uploads (host side, not capturable): run before capturing a batch.
hipError_t dxt_prepare();
hipError_t etc1_prepare();
hipError_t bc7_prepare();

// Persistent DXT1/DXT5/ETC1 kernels over all blocks of a format group,
// fed by the device counter (one fetch-and-add per workgroup and chunk).
hipError_t dxt_encode_batch(const SurfDesc *surfs, const uint32_t *block_surf, uint32_t total,
                            int dxt5, uint32_t *counter, const LaunchTune &tune,
                            hipStream_t stream);
hipError_t etc1_encode_batch(const SurfDesc *surfs, const uint32_t *block_surf, uint32_t total,
                             uint32_t *counter, const LaunchTune &tune, hipStream_t stream);
// BC7 group: gather blocks [g0, g0 + n) into a block array (n x 16 pixels)
// and scatter encoded blocks back to their surfaces.
hipError_t bc7_gather(const SurfDesc *surfs, const uint32_t *block_surf, uint32_t g0,
                      uint32_t n, uint32_t *blocks, hipStream_t stream);
hipError_t bc7_scatter(const SurfDesc *surfs, const uint32_t *block_surf, uint32_t g0,
                       uint32_t n, const uint8_t *enc, hipStream_t stream);

// Decoders and error reductions (decode.hip).  texture_decode writes the
// width x height RGBA8 image of an encoded DXT1/DXT5/ETC1/BPTC surface.
// block_error compares `original` with the decoded `encoded` surface
// (format >= 0) or with the RGBA8 image `encoded` (format kFormatRgba8):
// per-block premultiplied errors (block_err, optional) and their 128-bit
// total in sse[0] (low) and sse[1] (high), both overwritten.
constexpr int kFormatRgba8 = -1;
// BPTC decoded by the BC7 specification (FH_FORMAT_BPTC | FH_DECODE_SPEC).
constexpr int kFormatBptcSpec = 4;

// Mip levels and padding (mips.hip): dst = max(1, w/2) x max(1, h/2) box-filtered
// texels; pad4 copies w x h into ceil4(w) x ceil4(h) with transparent black.
hipError_t mip_level(const uint32_t *src, uint32_t w, uint32_t h, uint32_t *dst,
                     hipStream_t stream);
hipError_t pad4(const uint32_t *src, uint32_t w, uint32_t h, uint32_t *dst, hipStream_t stream);
hipError_t texture_decode(uint32_t format, const uint8_t *in, uint32_t *out, uint32_t width,
                          uint32_t height, hipStream_t stream);
hipError_t block_error(int format, const uint32_t *original, const void *encoded, uint32_t width,
                       uint32_t height, unsigned long long *block_err, unsigned long long *sse,
                       hipStream_t stream);

// BC7 candidate ranker (mlp.hip): 64 -> 256 -> 256 -> 512 MLP, fp32 master
// weights with Adam moments, bf16 copies for the matrix cores.
size_t ranker_num_params();

struct Ranker {
  int device = 0;
  float *master = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
  uint16_t *wbf = nullptr;
  uint16_t *w1_io = nullptr, *w2_io = nullptr, *w3_io = nullptr;  // [in][out]
  uint16_t *w1_oi = nullptr, *w2_oi = nullptr, *w3_oi = nullptr;  // [out][in]
  void *work = nullptr;
  size_t work_size = 0;
  bool work_training = false;
  int step = 0;
  // views into work, valid after forward()
  uint16_t *X = nullptr, *XT = nullptr, *h1 = nullptr, *h1T = nullptr, *h2 = nullptr,
           *h2T = nullptr, *dl = nullptr, *dlT = nullptr, *d2 = nullptr, *d2T = nullptr,
           *d1T = nullptr;
  float *logits = nullptr, *loss = nullptr;

  hipError_t init(int dev);
  void release();
  hipError_t refresh_bf16(hipStream_t s);
  // workspace for `rows` rows (a multiple of 32)
  hipError_t reserve(uint32_t rows, bool training);
  // logits of blocks first_block.. of r (n real rows, `rows` padded);
  // idx, if given, maps row -> block index relative to r.first_block
  hipError_t forward(const BlockRange &r, const uint32_t *idx, uint32_t n, uint32_t rows,
                     bool training, hipStream_t s);
  // gradients into grad (labels indexed like the rows, through idx if given)
  hipError_t backward(const int32_t *labels, const uint32_t *idx, uint32_t n, uint32_t rows,
                      hipStream_t s);
  hipError_t adam(float lr, hipStream_t s);
};

}  // namespace fh

// The C-ABI handle of a ranker (include/fasthip.h).
struct fh_ranker {
  fh::Ranker net;
  std::mutex mu;
};

namespace fh {


// Host-side table builders (tables.cpp); both mirror generator programs the
// reference documents.
void build_bc7_mode5_table(uint8_t out[256][2]);
void build_bc7_dxt6_table(uint8_t out[256][2][3]);
void build_bc7_single_color_table(uint32_t *out);  // kSingleColorEntries entries



// Sets the message fh_last_error returns on this thread (capi.cpp).
void set_error(const char *msg);

// Kernel timing (timing.cpp).  timing_begin records a start event on the
// stream when timing is enabled and returns its slot (-1 otherwise);
// timing_end records the matching stop event for kernel `tag`.
enum KernelTag {
  kTagBc7Scan,
  kTagBc7Select,
  kTagBc7Setup,
  kTagBc7Anneal,
  kTagBc7Pack,
  kTagBc7Guard,
  kTagDxt,
  kTagEtc1,
  kTagBc7Labels,
  kTagBc7Rank,
  kTagDecode,
  kTagBlockError,
  kNumKernelTags
};
int timing_begin(hipStream_t stream);
void timing_end(int slot, int tag, hipStream_t stream);

}  // namespace fh
