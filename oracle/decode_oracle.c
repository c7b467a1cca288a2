/* CPU restatement of FasTC's DXT1/DXT5 and ETC1 decoders and of the per-block
 * premultiplied error behind Image::ComputePSNR.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP decoders and error
 * reductions; nothing in the product links it.
 *
 * Pins (tests/test_decode_oracle.py):
 *   ora_dxt_decompress   hand-computed known-answer blocks (both colour modes,
 *                        both DXT5 alpha modes), the 565 -> 888 expansion
 *                        against round(v * 255 / (2^bits - 1)) for every value,
 *                        and decode(stb_dxt encode(solid colour)) through the
 *                        reference's own stb_dxt build in oracle/_ref.
 *   ora_etc1_decompress  the reference's rg_etc1::unpack_etc1_block, built from
 *                        the FasTC tree into oracle/_ref, on random blocks
 *                        (differential blocks whose second colour leaves
 *                        [0, 31] included) and on committed fixtures.
 *   ora_block_errors     brute force in float64 numpy against
 *                        Image::ComputePSNR's per-pixel formula.
 */
#include <stdint.h>
#include <string.h>

/* 565 channel expansion: Pixel::FromBits reads the byte-swapped word
 * (DXTEncoder/src/Decompressor.cpp:33-40, Base/src/Pixel.cpp:33-79), then
 * Pixel::ChangeBitDepth widens each channel by bit replication
 * (Base/src/Pixel.cpp:121-147, Base/include/FasTC/Bits.h:56-77). */
static void expand565(uint32_t c, int rgb[3]) {
  const int r = (int)((c >> 11) & 31), g = (int)((c >> 5) & 63), b = (int)(c & 31);
  rgb[0] = (r << 3) | (r >> 2);
  rgb[1] = (g << 2) | (g >> 4);
  rgb[2] = (b << 3) | (b >> 2);
}

static uint32_t pack_rgb(const int rgb[3]) {
  return (uint32_t)rgb[0] | ((uint32_t)rgb[1] << 8) | ((uint32_t)rgb[2] << 16);
}

/* DecompressDXT1Block (DXTEncoder/src/Decompressor.cpp:28-65).  `pixels`
 * keeps its alpha byte; the colours replace the low 24 bits.  check_order is
 * false for the colour half of DXT5, which always interpolates 4 colours. */
static void dxt1_block(const uint8_t *blk, uint32_t pixels[16], int check_order) {
  const uint32_t c0 = (uint32_t)blk[0] | ((uint32_t)blk[1] << 8);
  const uint32_t c1 = (uint32_t)blk[2] | ((uint32_t)blk[3] << 8);
  const uint32_t mod = (uint32_t)blk[4] | ((uint32_t)blk[5] << 8) | ((uint32_t)blk[6] << 16) |
                       ((uint32_t)blk[7] << 24);
  int a[3], b[3], c[3], d[3];
  expand565(c0, a);
  expand565(c1, b);
  for (int k = 0; k < 3; k++) {
    if (!check_order || c0 > c1) {
      c[k] = (a[k] * 2 + b[k]) / 3;
      d[k] = (a[k] + b[k] * 2) / 3;
    } else {
      /* three-colour mode: the fourth colour stays the zero pixel; the alpha
       * byte of the output is not touched (:52-55, :62-63) */
      c[k] = (a[k] + b[k]) / 2;
      d[k] = 0;
    }
  }
  const uint32_t pal[4] = {pack_rgb(a), pack_rgb(b), pack_rgb(c), pack_rgb(d)};
  for (int i = 0; i < 16; i++)
    pixels[i] = (pixels[i] & 0xFF000000u) | pal[(mod >> (2 * i)) & 3];
}

/* DecompressDXT5Block (DXTEncoder/src/Decompressor.cpp:67-93). */
static void dxt5_alpha_block(const uint8_t *blk, uint32_t pixels[16]) {
  const int a0 = blk[0], a1 = blk[1];
  int pal[8];
  pal[0] = a0;
  pal[1] = a1;
  if (a0 > a1) {
    for (int i = 2; i < 8; i++) pal[i] = ((8 - i) * a0 + (i - 1) * a1) / 7;
  } else {
    for (int i = 2; i < 6; i++) pal[i] = ((6 - i) * a0 + (i - 1) * a1) / 5;
    pal[6] = 0;
    pal[7] = 255;
  }
  uint64_t mod = 0;
  for (int k = 7; k >= 2; k--) mod = (mod << 8) | blk[k];
  for (int i = 0; i < 16; i++)
    pixels[i] = (pixels[i] & 0x00FFFFFFu) | ((uint32_t)pal[(mod >> (3 * i)) & 7] << 24);
}

/* DXTC::DecompressDXT1 / DecompressDXT5 (DXTEncoder/src/Decompressor.cpp:96-155):
 * block raster order, edge blocks clipped to the image.  The block buffer
 * starts as 0xFF bytes and is reused, so DXT1 output alpha is always 255. */
void ora_dxt_decompress(const uint8_t *in, uint32_t width, uint32_t height, int dxt5,
                        uint32_t *out) {
  const uint32_t bw = (width + 3) >> 2, bh = (height + 3) >> 2;
  const uint32_t bs = dxt5 ? 16 : 8;
  uint32_t block[16];
  memset(block, 0xFF, sizeof(block));
  for (uint32_t j = 0; j < bh; j++)
    for (uint32_t i = 0; i < bw; i++) {
      const uint8_t *blk = in + (size_t)(j * bw + i) * bs;
      if (dxt5) {
        dxt5_alpha_block(blk, block);
        dxt1_block(blk + 8, block, 0);
      } else {
        dxt1_block(blk, block, 1);
      }
      const uint32_t dw = width - i * 4 < 4 ? width - i * 4 : 4;
      const uint32_t dh = height - j * 4 < 4 ? height - j * 4 : 4;
      for (uint32_t y = 0; y < dh; y++)
        for (uint32_t x = 0; x < dw; x++)
          out[(size_t)(j * 4 + y) * width + i * 4 + x] = block[y * 4 + x];
    }
}

/* rg_etc1 tables (ETCEncoder/src/rg_etc1.cpp:371, :377). */
static const int kEtcInten[8][4] = {{-8, -2, 2, 8},       {-17, -5, 5, 17},   {-29, -9, 9, 29},
                                    {-42, -13, 13, 42},   {-60, -18, 18, 60}, {-80, -24, 24, 80},
                                    {-106, -33, 33, 106}, {-183, -47, 47, 183}};
static const int kEtcSelToIdx[4] = {2, 3, 1, 0};

static int clamp_int(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* rg_etc1::unpack_etc1_block with preserve_alpha = false
 * (ETCEncoder/src/rg_etc1.cpp:1147-1270): base colours from the 555+delta3
 * (:971-998, clamped to [0, 31] when the sum leaves it) or 444 (:1052-1066)
 * fields, the four intensity-offset colours clamped to [0, 255]
 * (:1076-1145, color_quad_u8::set), selectors from the LSB/MSB planes
 * (:598-610), alpha 255. */
static void etc1_block(const uint8_t *b, uint32_t px[16]) {
  const int diff = (b[3] >> 1) & 1, flip = b[3] & 1;
  const int tab[2] = {(b[3] >> 5) & 7, (b[3] >> 2) & 7};
  int base[2][3];
  if (diff) {
    for (int k = 0; k < 3; k++) {
      const int c5 = b[k] >> 3;
      int dc = b[k] & 7;
      if (dc >= 4) dc -= 8;
      const int c5b = clamp_int(c5 + dc, 0, 31);
      base[0][k] = (c5 << 3) | (c5 >> 2);
      base[1][k] = (c5b << 3) | (c5b >> 2);
    }
  } else {
    for (int k = 0; k < 3; k++) {
      const int c0 = b[k] >> 4, c1 = b[k] & 15;
      base[0][k] = (c0 << 4) | c0;
      base[1][k] = (c1 << 4) | c1;
    }
  }
  uint32_t col[2][4];
  for (int s = 0; s < 2; s++)
    for (int i = 0; i < 4; i++) {
      uint32_t v = 0xFF000000u;
      for (int k = 0; k < 3; k++)
        v |= (uint32_t)clamp_int(base[s][k] + kEtcInten[tab[s]][i], 0, 255) << (8 * k);
      col[s][i] = v;
    }
  for (int y = 0; y < 4; y++)
    for (int x = 0; x < 4; x++) {
      const int bit = x * 4 + y;
      const uint8_t *p = &b[7 - (bit >> 3)];
      const int lsb = (p[0] >> (bit & 7)) & 1, msb = (p[-2] >> (bit & 7)) & 1;
      const int sub = flip ? (y >= 2) : (x >= 2);
      px[y * 4 + x] = col[sub][kEtcSelToIdx[lsb | (msb << 1)]];
    }
}

/* ETCC::Decompress (ETCEncoder/src/Decompressor.cpp:27-48). */
void ora_etc1_decompress(const uint8_t *in, uint32_t width, uint32_t height, uint32_t *out) {
  const uint32_t bw = (width + 3) / 4, bh = (height + 3) / 4;
  for (uint32_t j = 0; j < bh; j++)
    for (uint32_t i = 0; i < bw; i++) {
      uint32_t px[16];
      etc1_block(in + (size_t)(j * bw + i) * 8, px);
      const uint32_t dw = width - i * 4 < 4 ? width - i * 4 : 4;
      const uint32_t dh = height - j * 4 < 4 ? height - j * 4 : 4;
      for (uint32_t y = 0; y < dh; y++)
        for (uint32_t x = 0; x < dw; x++)
          out[(size_t)(j * 4 + y) * width + i * 4 + x] = px[y * 4 + x];
    }
}

/* Per-block premultiplied RGB error over the pixels inside the image, in
 * units of 1/255^2 of Image::ComputePSNR's squared differences
 * (Base/src/Image.cpp:226-245): sum of (a*c - a'*c')^2.  Blocks in raster
 * order; returns the image total. */
uint64_t ora_block_errors(const uint32_t *a, const uint32_t *b, uint32_t width, uint32_t height,
                          uint64_t *block_err) {
  const uint32_t bw = (width + 3) / 4, bh = (height + 3) / 4;
  uint64_t total = 0;
  for (uint32_t j = 0; j < bh; j++)
    for (uint32_t i = 0; i < bw; i++) {
      uint64_t e = 0;
      for (uint32_t y = j * 4; y < j * 4 + 4 && y < height; y++)
        for (uint32_t x = i * 4; x < i * 4 + 4 && x < width; x++) {
          const uint32_t p = a[(size_t)y * width + x], q = b[(size_t)y * width + x];
          const int64_t pa = p >> 24, qa = q >> 24;
          for (int c = 0; c < 3; c++) {
            const int64_t d = pa * (int64_t)((p >> (8 * c)) & 0xFF) -
                              qa * (int64_t)((q >> (8 * c)) & 0xFF);
            e += (uint64_t)(d * d);
          }
        }
      if (block_err) block_err[j * bw + i] = e;
      total += e;
    }
  return total;
}
