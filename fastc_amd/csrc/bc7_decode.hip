// BC7 (BPTC) decoder for gfx950: BPTCC::Decompress
// (BPTCEncoder/src/Decompressor.cpp:32-367), one lane per block.
//
// Like the reference, mode 4 blocks with index mode 1 interpolate with the
// per-mode index widths rather than the swapped ones (Decompressor.cpp:263-295),
// so decoded images match the reference decoder byte for byte.  Blocks with
// the reserved mode (no mode bit in the first byte) decode to zero.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bc7_block_decode.h"
#include "internal.h"

namespace fh {
namespace bc7dec {

__global__ void decode_kernel(const uint8_t *in, uint32_t *out, uint32_t width, uint32_t height) {
  const uint32_t bw = (width + 3) / 4, bh = (height + 3) / 4;
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= bw * bh) return;
  uint32_t px[16];
  decode_block(reinterpret_cast<const uint4 *>(in)[b], px);
  const uint32_t bx = b % bw, by = b / bw;
  for (int y = 0; y < 4; y++) {
    const uint32_t yy = by * 4 + y;
    if (yy >= height) break;
    for (int x = 0; x < 4; x++) {
      const uint32_t xx = bx * 4 + x;
      if (xx < width) out[(size_t)yy * width + xx] = px[y * 4 + x];
    }
  }
}

}  // namespace bc7dec

hipError_t bc7_decode(const uint8_t *in, uint32_t *out, uint32_t width, uint32_t height,
                      hipStream_t stream) {
  const uint32_t nb = ((width + 3) / 4) * ((height + 3) / 4);
  if (nb == 0) return hipSuccess;
  const int t = timing_begin(stream);
  hipLaunchKernelGGL(bc7dec::decode_kernel, dim3((nb + 255) / 256), dim3(256), 0, stream, in,
                     out, width, height);
  timing_end(t, kTagBc7Decode, stream);
  return hipGetLastError();
}

}  // namespace fh
