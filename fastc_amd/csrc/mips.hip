// Mip-chain generation and 4-texel padding for gfx950.
//
// FasTC encodes the levels it is given; a texture batch needs the levels
// themselves.  A level is the rounded 2x2 box filter of the one above it
// (2x1 or 1x2 once a side is 1): per channel (a + b + c + d + 2) >> 2, or
// (a + b + 1) >> 1.  Sides halve rounding down (an odd last row or column
// is dropped).  Padding copies a level into a grid of whole 4x4 blocks
// filled with transparent black, as CompressImage pads images whose sides
// are not multiples of 4 (Core/src/TexComp.cpp:382-393).  Both kernels are
// one thread per output texel and HBM-bound.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.h"

namespace fh {
namespace mips {

// Per-byte rounded mean of two or four RGBA8 texels, computed on the even and
// odd bytes as 16-bit lanes of 32-bit words (sums stay below 2^10).
__device__ __forceinline__ uint32_t mean4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
  const uint32_t lo = (a & 0x00FF00FFu) + (b & 0x00FF00FFu) + (c & 0x00FF00FFu) +
                      (d & 0x00FF00FFu) + 0x00020002u;
  const uint32_t hi = ((a >> 8) & 0x00FF00FFu) + ((b >> 8) & 0x00FF00FFu) +
                      ((c >> 8) & 0x00FF00FFu) + ((d >> 8) & 0x00FF00FFu) + 0x00020002u;
  return ((lo >> 2) & 0x00FF00FFu) | (((hi >> 2) & 0x00FF00FFu) << 8);
}

__device__ __forceinline__ uint32_t mean2(uint32_t a, uint32_t b) {
  const uint32_t lo = (a & 0x00FF00FFu) + (b & 0x00FF00FFu) + 0x00010001u;
  const uint32_t hi = ((a >> 8) & 0x00FF00FFu) + ((b >> 8) & 0x00FF00FFu) + 0x00010001u;
  return ((lo >> 1) & 0x00FF00FFu) | (((hi >> 1) & 0x00FF00FFu) << 8);
}

__global__ void __launch_bounds__(256) level_kernel(const uint32_t *src, uint32_t w, uint32_t h,
                                                    uint32_t *dst, uint32_t dw, uint32_t dh) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)dw * dh) return;
  const uint32_t x = (uint32_t)(t % dw), y = (uint32_t)(t / dw);
  uint32_t v;
  if (w > 1 && h > 1) {
    const uint32_t *r0 = src + (size_t)(2 * y) * w + 2 * x;
    const uint32_t *r1 = r0 + w;
    v = mean4(r0[0], r1[0], r0[1], r1[1]);
  } else if (w > 1) {
    v = mean2(src[2 * x], src[2 * x + 1]);
  } else {
    v = mean2(src[(size_t)(2 * y) * w], src[(size_t)(2 * y + 1) * w]);
  }
  dst[(size_t)y * dw + x] = v;
}

__global__ void __launch_bounds__(256) pad_kernel(const uint32_t *src, uint32_t w, uint32_t h,
                                                  uint32_t *dst, uint32_t pw, uint32_t ph) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)pw * ph) return;
  const uint32_t x = (uint32_t)(t % pw), y = (uint32_t)(t / pw);
  dst[(size_t)y * pw + x] = (x < w && y < h) ? src[(size_t)y * w + x] : 0u;
}

}  // namespace mips

hipError_t mip_level(const uint32_t *src, uint32_t w, uint32_t h, uint32_t *dst,
                     hipStream_t stream) {
  const uint32_t dw = w > 1 ? w / 2 : 1, dh = h > 1 ? h / 2 : 1;
  const uint64_t n = (uint64_t)dw * dh;
  hipLaunchKernelGGL(mips::level_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream,
                     src, w, h, dst, dw, dh);
  return hipGetLastError();
}

hipError_t pad4(const uint32_t *src, uint32_t w, uint32_t h, uint32_t *dst, hipStream_t stream) {
  const uint32_t pw = (w + 3) & ~3u, ph = (h + 3) & ~3u;
  const uint64_t n = (uint64_t)pw * ph;
  hipLaunchKernelGGL(mips::pad_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream,
                     src, w, h, dst, pw, ph);
  return hipGetLastError();
}

}  // namespace fh
