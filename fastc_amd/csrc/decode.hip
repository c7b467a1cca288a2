// Texture decoders and encoded-image error reductions for gfx950.
//
// Decoders, one lane per 4x4 block, bit-exact with the reference's:
//   DXT1/DXT5  DXTC::DecompressDXT1/DXT5 (DXTEncoder/src/Decompressor.cpp:28-155):
//              565 endpoints widened by bit replication (Base/src/Pixel.cpp:121-147),
//              interpolants truncated, three-colour mode only for DXT1 with
//              c0 <= c1, and DXT1 output alpha always 255 (the reference keeps the
//              block buffer's 0xFF alpha bytes);
//   ETC1       rg_etc1::unpack_etc1_block with preserve_alpha = false
//              (ETCEncoder/src/rg_etc1.cpp:1147-1270), the differential base
//              clamped to [0, 31] when base + delta leaves it;
//   BPTC       bc7dec::decode_block (bc7_block_decode.h).
// Edge blocks are clipped to the image as in the reference's raster loops.
//
// block_error fuses decoding with the comparison behind Image::ComputePSNR
// (Base/src/Image.cpp:205-255): per pixel sum over RGB of (a*c - a'*c')^2,
// which is 255^2 times the reference's squared difference of (a/255)*c.  Every
// term is an integer below 2^32, so per-block sums are exact and the image
// total (a 128-bit integer, built from 64-bit atomics with carries) does not
// depend on the order of the reduction.  The decoded image is never written.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fasthip.h"
#include "bc7_block_decode.h"
#include "internal.h"

namespace fh {
namespace dec {

// Small and large modifiers of rg_etc1's g_etc1_inten_tables (rg_etc1.cpp:371):
// row t is {-large, -small, small, large}.
__constant__ int kEtcSmall[8] = {2, 5, 9, 13, 18, 24, 33, 47};
__constant__ int kEtcLarge[8] = {8, 17, 29, 42, 60, 80, 106, 183};

__device__ __forceinline__ uint32_t expand565(uint32_t c) {
  const uint32_t r = (c >> 11) & 31u, g = (c >> 5) & 63u, b = c & 31u;
  return ((r << 3) | (r >> 2)) | (((g << 2) | (g >> 4)) << 8) | (((b << 3) | (b >> 2)) << 16);
}

// Colour half (DecompressDXT1Block): replaces the low 24 bits of px.
__device__ __forceinline__ void dxt_colour(uint32_t ends, uint32_t idx, bool check_order,
                                           uint32_t px[16]) {
  const uint32_t c0 = ends & 0xFFFFu, c1 = ends >> 16;
  const uint32_t e0 = expand565(c0), e1 = expand565(c1);
  const bool four = !check_order || c0 > c1;
  uint32_t p2 = 0, p3 = 0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const uint32_t a = (e0 >> (8 * k)) & 0xFFu, b = (e1 >> (8 * k)) & 0xFFu;
    if (four) {
      p2 |= ((2 * a + b) / 3) << (8 * k);
      p3 |= ((a + 2 * b) / 3) << (8 * k);
    } else {
      p2 |= ((a + b) >> 1) << (8 * k);
    }
  }
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const uint32_t s = (idx >> (2 * i)) & 3u;
    const uint32_t c = s == 0 ? e0 : (s == 1 ? e1 : (s == 2 ? p2 : p3));
    px[i] = (px[i] & 0xFF000000u) | c;
  }
}

// Alpha half of DXT5 (DecompressDXT5Block): replaces the top byte of px.
__device__ __forceinline__ void dxt5_alpha(uint32_t lo, uint32_t hi, uint32_t px[16]) {
  const uint32_t a0 = lo & 0xFFu, a1 = (lo >> 8) & 0xFFu;
  const uint64_t bits = (uint64_t)(lo >> 16) | ((uint64_t)hi << 16);
  const bool eight = a0 > a1;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const uint32_t s = (uint32_t)(bits >> (3 * i)) & 7u;
    uint32_t a;
    if (s == 0) {
      a = a0;
    } else if (s == 1) {
      a = a1;
    } else if (eight) {
      a = ((8 - s) * a0 + (s - 1) * a1) / 7;
    } else if (s < 6) {
      a = ((6 - s) * a0 + (s - 1) * a1) / 5;
    } else {
      a = s == 6 ? 0u : 255u;
    }
    px[i] = (px[i] & 0x00FFFFFFu) | (a << 24);
  }
}

__device__ __forceinline__ int clamp_int(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// rg_etc1::unpack_etc1_block.  Bytes b0..b7 of the block are lo (b0 lowest)
// and hi (b4 lowest).
__device__ __forceinline__ void etc1_block(uint32_t lo, uint32_t hi, uint32_t px[16]) {
  const uint32_t b3 = lo >> 24;
  const bool diff = (b3 >> 1) & 1u, flip = b3 & 1u;
  const uint32_t t0 = (b3 >> 5) & 7u, t1 = (b3 >> 2) & 7u;
  int base0[3], base1[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int byte = (int)((lo >> (8 * k)) & 0xFFu);
    if (diff) {
      const int c5 = byte >> 3;
      const int dc = ((byte & 7) ^ 4) - 4;  // sign-extended 3-bit delta
      const int c5b = clamp_int(c5 + dc, 0, 31);
      base0[k] = (c5 << 3) | (c5 >> 2);
      base1[k] = (c5b << 3) | (c5b >> 2);
    } else {
      const int c0 = byte >> 4, c1 = byte & 15;
      base0[k] = (c0 << 4) | c0;
      base1[k] = (c1 << 4) | c1;
    }
  }
  const int s0 = kEtcSmall[t0], l0 = kEtcLarge[t0], s1 = kEtcSmall[t1], l1 = kEtcLarge[t1];
  // selector planes: bit x*4+y of the LSB word (b6:b7) and the MSB word (b4:b5)
  const uint32_t lsb = ((hi >> 8) & 0xFF00u) | (hi >> 24);
  const uint32_t msb = ((hi & 0xFFu) << 8) | ((hi >> 8) & 0xFFu);
#pragma unroll
  for (int y = 0; y < 4; y++) {
#pragma unroll
    for (int x = 0; x < 4; x++) {
      const int bit = x * 4 + y;
      const bool second = flip ? (y >= 2) : (x >= 2);
      const uint32_t l = (lsb >> bit) & 1u, m = (msb >> bit) & 1u;
      // ETC1 selector value m:l = 0 -> +small, 1 -> +large, 2 -> -small, 3 -> -large
      const int mag = l ? (second ? l1 : l0) : (second ? s1 : s0);
      const int mod = m ? -mag : mag;
      uint32_t v = 0xFF000000u;
#pragma unroll
      for (int k = 0; k < 3; k++)
        v |= (uint32_t)clamp_int((second ? base1[k] : base0[k]) + mod, 0, 255) << (8 * k);
      px[y * 4 + x] = v;
    }
  }
}

template <int FMT>
__device__ __forceinline__ void decode_one(const uint8_t *in, uint32_t b, uint32_t px[16]) {
  if constexpr (FMT == FH_FORMAT_DXT1) {
    const uint2 v = reinterpret_cast<const uint2 *>(in)[b];
#pragma unroll
    for (int i = 0; i < 16; i++) px[i] = 0xFF000000u;
    dxt_colour(v.x, v.y, true, px);
  } else if constexpr (FMT == FH_FORMAT_DXT5) {
    const uint4 v = reinterpret_cast<const uint4 *>(in)[b];
#pragma unroll
    for (int i = 0; i < 16; i++) px[i] = 0;
    dxt5_alpha(v.x, v.y, px);
    dxt_colour(v.z, v.w, false, px);
  } else if constexpr (FMT == FH_FORMAT_ETC1) {
    const uint2 v = reinterpret_cast<const uint2 *>(in)[b];
    etc1_block(v.x, v.y, px);
  } else if constexpr (FMT == kFormatBptcSpec) {
    bc7dec::decode_block<true>(reinterpret_cast<const uint4 *>(in)[b], px);
  } else {
    bc7dec::decode_block(reinterpret_cast<const uint4 *>(in)[b], px);
  }
}

// VEC: width % 4 == 0 and 16-byte aligned rows, so full block rows move as
// one 16-byte access.
template <int FMT, bool VEC>
__global__ void __launch_bounds__(256)
    decode_kernel(const uint8_t *in, uint32_t *out, uint32_t width, uint32_t height) {
  const uint32_t bw = (width + 3) / 4, bh = (height + 3) / 4;
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= bw * bh) return;
  uint32_t px[16];
  decode_one<FMT>(in, b, px);
  const uint32_t bx = b % bw, by = b / bw;
  const bool full = VEC && bx * 4 + 4 <= width;
#pragma unroll
  for (int y = 0; y < 4; y++) {
    const uint32_t yy = by * 4 + y;
    if (yy >= height) break;
    uint32_t *row = out + (size_t)yy * width + bx * 4;
    if (full) {
      *reinterpret_cast<uint4 *>(row) = make_uint4(px[4 * y], px[4 * y + 1], px[4 * y + 2],
                                                   px[4 * y + 3]);
    } else {
#pragma unroll
      for (int x = 0; x < 4; x++)
        if (bx * 4 + x < width) row[x] = px[4 * y + x];
    }
  }
}

// Exact per-pixel error: each square is at most 65025^2 < 2^32, but the three
// together can exceed 2^32, so they are added in 64 bits.
__device__ __forceinline__ uint64_t pixel_error64(uint32_t p, uint32_t q) {
  const int pa = (int)(p >> 24), qa = (int)(q >> 24);
  uint64_t e = 0;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const int d = pa * (int)((p >> (8 * c)) & 0xFFu) - qa * (int)((q >> (8 * c)) & 0xFFu);
    const uint32_t ad = (uint32_t)(d < 0 ? -d : d);  // squared in unsigned arithmetic
    e += (uint64_t)(ad * ad);
  }
  return e;
}

template <int FMT, bool VEC>
__global__ void __launch_bounds__(256)
    error_kernel(const uint32_t *orig, const void *enc, uint32_t width, uint32_t height,
                 unsigned long long *block_err, unsigned long long *sse) {
  const uint32_t bw = (width + 3) / 4, bh = (height + 3) / 4;
  uint64_t err = 0;
  // persistent: a few workgroups per CU stride over the blocks, so that the
  // total needs only one atomic accumulation per workgroup
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < bw * bh;
       b += gridDim.x * blockDim.x) {
    uint64_t be = 0;
    const uint32_t bx = b % bw, by = b / bw;
    const bool full = VEC && bx * 4 + 4 <= width;
    uint32_t px[16];
    if constexpr (FMT != kFormatRgba8) decode_one<FMT>(static_cast<const uint8_t *>(enc), b, px);
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const uint32_t yy = by * 4 + y;
      if (yy >= height) break;
      const size_t off = (size_t)yy * width + bx * 4;
      if (full) {
        const uint4 o = *reinterpret_cast<const uint4 *>(orig + off);
        uint4 d;
        if constexpr (FMT == kFormatRgba8)
          d = *reinterpret_cast<const uint4 *>(static_cast<const uint32_t *>(enc) + off);
        else
          d = make_uint4(px[4 * y], px[4 * y + 1], px[4 * y + 2], px[4 * y + 3]);
        be += pixel_error64(o.x, d.x) + pixel_error64(o.y, d.y) + pixel_error64(o.z, d.z) +
              pixel_error64(o.w, d.w);
      } else {
#pragma unroll
        for (int x = 0; x < 4; x++) {
          if (bx * 4 + x >= width) continue;
          uint32_t d;
          if constexpr (FMT == kFormatRgba8)
            d = static_cast<const uint32_t *>(enc)[off + x];
          else
            d = px[4 * y + x];
          be += pixel_error64(orig[off + x], d);
        }
      }
    }
    if (block_err) block_err[b] = be;
    err += be;
  }
  // workgroup sum, then one 128-bit accumulation per workgroup
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) err += __shfl_xor(err, o);
  __shared__ uint64_t part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = err;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint64_t s = part[0] + part[1] + part[2] + part[3];
    if (s) {
      const unsigned long long old = atomicAdd(&sse[0], (unsigned long long)s);
      if (old + s < old) atomicAdd(&sse[1], 1ull);
    }
  }
}

// Compute units of the current device (cached per device).
int num_cus() {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cus[dev]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        v <= 0)
      v = 256;
    cus[dev] = v;
  }
  return cus[dev];
}

template <int FMT>
hipError_t launch_decode(const uint8_t *in, uint32_t *out, uint32_t w, uint32_t h,
                         hipStream_t stream) {
  const uint32_t nb = ((w + 3) / 4) * ((h + 3) / 4);
  const bool vec = (w % 4) == 0 && (reinterpret_cast<uintptr_t>(out) & 15) == 0;
  if (vec)
    hipLaunchKernelGGL((decode_kernel<FMT, true>), dim3((nb + 255) / 256), dim3(256), 0, stream,
                       in, out, w, h);
  else
    hipLaunchKernelGGL((decode_kernel<FMT, false>), dim3((nb + 255) / 256), dim3(256), 0,
                       stream, in, out, w, h);
  return hipGetLastError();
}

template <int FMT>
hipError_t launch_error(const uint32_t *orig, const void *enc, uint32_t w, uint32_t h,
                        unsigned long long *block_err, unsigned long long *sse,
                        hipStream_t stream) {
  const uint32_t nb = ((w + 3) / 4) * ((h + 3) / 4);
  bool vec = (w % 4) == 0 && (reinterpret_cast<uintptr_t>(orig) & 15) == 0;
  if (FMT == kFormatRgba8) vec = vec && (reinterpret_cast<uintptr_t>(enc) & 15) == 0;
  uint32_t grid = (nb + 255) / 256;
  const uint32_t cap = (uint32_t)num_cus() * 4u;
  if (grid > cap) grid = cap;
  if (vec)
    hipLaunchKernelGGL((error_kernel<FMT, true>), dim3(grid), dim3(256), 0, stream, orig, enc, w,
                       h, block_err, sse);
  else
    hipLaunchKernelGGL((error_kernel<FMT, false>), dim3(grid), dim3(256), 0, stream, orig, enc,
                       w, h, block_err, sse);
  return hipGetLastError();
}

}  // namespace dec

hipError_t texture_decode(uint32_t format, const uint8_t *in, uint32_t *out, uint32_t width,
                          uint32_t height, hipStream_t stream) {
  if (width == 0 || height == 0) return hipSuccess;
  const int t = timing_begin(stream);
  hipError_t e = hipErrorInvalidValue;
  switch (format) {
    case FH_FORMAT_DXT1: e = dec::launch_decode<FH_FORMAT_DXT1>(in, out, width, height, stream); break;
    case FH_FORMAT_DXT5: e = dec::launch_decode<FH_FORMAT_DXT5>(in, out, width, height, stream); break;
    case FH_FORMAT_ETC1: e = dec::launch_decode<FH_FORMAT_ETC1>(in, out, width, height, stream); break;
    case FH_FORMAT_BPTC: e = dec::launch_decode<FH_FORMAT_BPTC>(in, out, width, height, stream); break;
    case kFormatBptcSpec: e = dec::launch_decode<kFormatBptcSpec>(in, out, width, height, stream); break;
  }
  timing_end(t, kTagDecode, stream);
  return e;
}

hipError_t block_error(int format, const uint32_t *original, const void *encoded, uint32_t width,
                       uint32_t height, unsigned long long *block_err, unsigned long long *sse,
                       hipStream_t stream) {
  hipError_t e = fill_u32(reinterpret_cast<uint32_t *>(sse), 4, 0, stream);
  if (e != hipSuccess || width == 0 || height == 0) return e;
  const int t = timing_begin(stream);
  e = hipErrorInvalidValue;
  switch (format) {
    case kFormatRgba8:
      e = dec::launch_error<kFormatRgba8>(original, encoded, width, height, block_err, sse, stream);
      break;
    case FH_FORMAT_DXT1:
      e = dec::launch_error<FH_FORMAT_DXT1>(original, encoded, width, height, block_err, sse, stream);
      break;
    case FH_FORMAT_DXT5:
      e = dec::launch_error<FH_FORMAT_DXT5>(original, encoded, width, height, block_err, sse, stream);
      break;
    case FH_FORMAT_ETC1:
      e = dec::launch_error<FH_FORMAT_ETC1>(original, encoded, width, height, block_err, sse, stream);
      break;
    case FH_FORMAT_BPTC:
      e = dec::launch_error<FH_FORMAT_BPTC>(original, encoded, width, height, block_err, sse, stream);
      break;
    case kFormatBptcSpec:
      e = dec::launch_error<kFormatBptcSpec>(original, encoded, width, height, block_err, sse, stream);
      break;
  }
  timing_end(t, kTagBlockError, stream);
  return e;
}

}  // namespace fh
