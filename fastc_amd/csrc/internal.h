// Internal interfaces between the C-ABI layer (capi.cpp) and the HIP kernel
// translation units.  Not installed; callers use include/fasthip.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

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

struct Bc7Params {
  uint32_t quality;      // annealing steps
  uint32_t block_modes;  // mask of allowed BC7 modes
  uint64_t seed;         // annealing stream key
  uint64_t block_id_base;  // added to the raster block index to key the stream
  uint32_t q0_guard;     // keep the quality-0 block where annealing decodes worse
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
// records, queues and annealing states).
size_t bc7_work_bytes(uint32_t num_blocks);

// Largest block count one bc7_encode call takes; callers split bigger jobs.
constexpr uint32_t kBc7MaxBlocksPerCall = 1u << 20;

hipError_t bc7_encode(const BlockRange &r, const Bc7Params &p, const Bc7Scratch &s, void *work,
                      hipStream_t stream);

// Exhaustive quality-0 search over all 285 (mode, slot) candidates of the
// candidate ranker: labels[i] = mode * 64 + slot of the smallest decoded
// premultiplied error (errors[i], optional; Image::ComputePSNR's per-pixel
// term times 255^2).  best: n words of scratch.
hipError_t bc7_exhaustive_labels(const BlockRange &r, unsigned long long *best, int32_t *labels,
                                 unsigned long long *errors, hipStream_t stream);

hipError_t dxt_encode(const BlockRange &r, int dxt5, hipStream_t stream);
hipError_t etc1_encode(const BlockRange &r, hipStream_t stream);

hipError_t bc7_decode(const uint8_t *in, uint32_t *out, uint32_t width, uint32_t height,
                      hipStream_t stream);

// Host-side table builders (tables.cpp); both mirror generator programs the
// reference documents.
void build_bc7_mode5_table(uint8_t out[256][2]);
void build_bc7_dxt6_table(uint8_t out[256][2][3]);
void build_bc7_single_color_table(uint32_t *out);  // kSingleColorEntries entries

constexpr int kSingleColorEntries = 5 * 5 * 4 * 256;

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
  kTagBc7Decode,
  kTagBc7Labels,
  kNumKernelTags
};
int timing_begin(hipStream_t stream);
void timing_end(int slot, int tag, hipStream_t stream);

}  // namespace fh
