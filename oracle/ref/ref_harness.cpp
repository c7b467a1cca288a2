// Harness that drives the reference's own third-party encoders, compiled from
// their sources in the FasTC tree (see oracle/ref/Makefile):
//   DXTEncoder/src/stb_dxt.h   (stb_dxt v1.06, DXT1/DXT5)
//   ETCEncoder/src/rg_etc1.cpp (rg_etc1, ETC1)
// The loops restate FasTC's wrappers around them:
//   DXTC::CompressImageDXT1/DXT5 (DXTEncoder/src/Compressor.cpp:47-95)
//   ETCC::Compress_RG            (ETCEncoder/src/Compressor.cpp:26-54)
//   ETCC::Decompress             (ETCEncoder/src/Decompressor.cpp:24-48)
// TEST INFRASTRUCTURE ONLY: used by tests/ and the benchmark's CPU baseline.
#include <stdint.h>
#include <string.h>

#define STB_DXT_IMPLEMENTATION
#include "stb_dxt.h"
#include "rg_etc1.h"

extern "C" {

// Full-image DXT1 (alpha = 0) or DXT5 (alpha = 1); blocks in raster order.
void ref_dxt_compress(const uint8_t *rgba, uint32_t width, uint32_t height, int alpha,
                      uint8_t *out) {
  uint8_t block[64];
  const uint32_t *px = reinterpret_cast<const uint32_t *>(rgba);
  for (uint32_t j = 0; j + 4 <= height; j += 4)
    for (uint32_t i = 0; i < width; i += 4) {
      for (int r = 0; r < 4; r++) memcpy(block + 16 * r, px + (j + r) * width + i, 16);
      stb_compress_dxt_block(out, block, alpha, STB_DXT_DITHER);
      out += alpha ? 16 : 8;
    }
}

// Full-image ETC1 at rg_etc1::cLowQuality without dithering, as FasTC's
// Compress_RG configures it (it calls pack_etc1_block_init() first).
void ref_etc1_compress(const uint8_t *rgba, uint32_t width, uint32_t height, uint8_t *out) {
  rg_etc1::etc1_pack_params params;
  params.m_quality = rg_etc1::cLowQuality;
  rg_etc1::pack_etc1_block_init();
  const uint32_t *px = reinterpret_cast<const uint32_t *>(rgba);
  for (uint32_t j = 0; j + 4 <= height; j += 4)
    for (uint32_t i = 0; i < width; i += 4) {
      uint32_t block[16];
      for (int r = 0; r < 4; r++) memcpy(block + 4 * r, px + (j + r) * width + i, 16);
      rg_etc1::pack_etc1_block(out, block, params);
      out += 8;
    }
}

void ref_etc1_decompress(const uint8_t *in, uint32_t width, uint32_t height, uint8_t *rgba) {
  uint32_t *out = reinterpret_cast<uint32_t *>(rgba);
  const uint32_t bw = (width + 3) / 4, bh = (height + 3) / 4;
  for (uint32_t j = 0; j < bh; j++)
    for (uint32_t i = 0; i < bw; i++) {
      uint32_t px[16];
      rg_etc1::unpack_etc1_block(in + (j * bw + i) * 8, px);
      for (uint32_t y = 0; y < 4 && j * 4 + y < height; y++)
        for (uint32_t x = 0; x < 4 && i * 4 + x < width; x++)
          out[(j * 4 + y) * width + i * 4 + x] = px[y * 4 + x];
    }
}

}  // extern "C"
