/* Seeded synthetic RGBA8 images: a C implementation of fastc_amd/synth.py's
 * generate() (same integer arithmetic, same bytes), for workload sizes where
 * numpy is too slow.  Host code only; built into fastc_amd/libsynth.so. */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static uint64_t sm64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static uint64_t hash2(uint64_t a, uint64_t b) {
  uint64_t h = sm64(0x5EED5EED5EED5EEDull);
  h = sm64(h ^ a);
  return sm64(h ^ b);
}
static uint64_t hash3(uint64_t a, uint64_t b, uint64_t c) { return sm64(hash2(a, b) ^ c); }
static uint64_t hash4(uint64_t a, uint64_t b, uint64_t c, uint64_t d) {
  return sm64(hash3(a, b, c) ^ d);
}
static uint64_t key(int64_t y, int64_t x) { return ((uint64_t)y << 32) | (uint64_t)x; }

static const int64_t kDirs[16][2] = {{64, 0},    {59, 24},   {45, 45},   {24, 59},
                                     {0, 64},    {-24, 59},  {-45, 45},  {-59, 24},
                                     {-64, 0},   {-59, -24}, {-45, -45}, {-24, -59},
                                     {0, -64},   {24, -59},  {45, -45},  {59, -24}};

/* Returns 0 on success, -1 on bad arguments or allocation failure. */
int synth_generate(uint32_t *out, int64_t width, int64_t height, uint64_t seed, int edges,
                   int64_t x0, int64_t y0) {
  if (width % 4 || height % 4 || x0 % 4 || y0 % 4 || width <= 0 || height <= 0) return -1;
  int64_t *acc = (int64_t *)calloc((size_t)(width * height), sizeof(int64_t));
  uint8_t *rgb = (uint8_t *)malloc((size_t)(width * height) * 3);
  if (!acc || !rgb) {
    free(acc);
    free(rgb);
    return -1;
  }
  for (int ch = 0; ch < 3; ch++) {
    memset(acc, 0, sizeof(int64_t) * (size_t)(width * height));
    for (int octave = 0; octave < 4; octave++) {
      const int shift = 5 - octave;
      const int64_t amp = 8 >> octave;
      const int64_t lx0 = x0 >> shift, ly0 = y0 >> shift;
      const int64_t nlx = ((x0 + width - 1) >> shift) - lx0 + 2;
      const int64_t nly = ((y0 + height - 1) >> shift) - ly0 + 2;
      int64_t *lat = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nlx * nly));
      if (!lat) {
        free(acc);
        free(rgb);
        return -1;
      }
      for (int64_t gy = 0; gy < nly; gy++)
        for (int64_t gx = 0; gx < nlx; gx++)
          lat[gy * nlx + gx] =
              (int64_t)(hash4(seed, 1 + ch, octave, key(ly0 + gy, lx0 + gx)) & 0xFF);
      for (int64_t y = 0; y < height; y++) {
        const int64_t yy = y0 + y;
        const int64_t ly = (yy >> shift) - ly0;
        const int64_t ty = (yy & ((1 << shift) - 1)) << (16 - shift);
        const int64_t sy = (ty * ty * (3 * 65536 - 2 * ty)) >> 32;
        for (int64_t x = 0; x < width; x++) {
          const int64_t xx = x0 + x;
          const int64_t lx = (xx >> shift) - lx0;
          const int64_t tx = (xx & ((1 << shift) - 1)) << (16 - shift);
          const int64_t sx = (tx * tx * (3 * 65536 - 2 * tx)) >> 32;
          const int64_t l00 = lat[ly * nlx + lx], l10 = lat[ly * nlx + lx + 1];
          const int64_t l01 = lat[(ly + 1) * nlx + lx], l11 = lat[(ly + 1) * nlx + lx + 1];
          const int64_t top = (l00 << 16) + (l10 - l00) * sx;
          const int64_t bot = (l01 << 16) + (l11 - l01) * sx;
          acc[y * width + x] += amp * ((top << 16) + (bot - top) * sy);
        }
      }
      free(lat);
    }
    for (int64_t y = 0; y < height; y++)
      for (int64_t x = 0; x < width; x++) {
        const uint64_t nz = hash3(seed, 10, key(y0 + y, x0 + x));
        int64_t v = acc[y * width + x] / ((int64_t)15 << 32);
        v += (int64_t)(((nz >> (16 * ch)) & 0xFFFF) % 17) - 8;
        v = v < 0 ? 0 : (v > 255 ? 255 : v);
        rgb[(y * width + x) * 3 + ch] = (uint8_t)v;
      }
  }
  free(acc);
  for (int64_t by = 0; by < height / 4; by++)
    for (int64_t bx = 0; bx < width / 4; bx++) {
      const uint64_t bk = key(y0 / 4 + by, x0 / 4 + bx);
      const uint64_t hb1 = hash3(seed, 20, bk), hb2 = hash3(seed, 30, bk);
      const uint64_t fcol = hash3(seed, 31, bk);
      const int64_t aroll = (int64_t)(hb1 % 100);
      const int64_t adir = (int64_t)((hb1 >> 8) % 16);
      const int64_t dx = kDirs[adir][0], dy = kDirs[adir][1];
      int64_t pmin = INT64_MAX, pmax = INT64_MIN;
      for (int cx = -3; cx <= 3; cx += 6)
        for (int cy = -3; cy <= 3; cy += 6) {
          const int64_t c = dx * cx + dy * cy;
          pmin = c < pmin ? c : pmin;
          pmax = c > pmax ? c : pmax;
        }
      const int flat = (int64_t)(hb2 % 100) < 10;
      int edge = 0;
      uint64_t he = 0, c0 = 0, c1 = 0;
      if (edges) {
        he = hash3(seed, 40, bk);
        edge = (int64_t)(he % 100) < 5 && !flat;
        if (edge) {
          c0 = hash3(seed, 41, bk);
          c1 = hash3(seed, 42, bk);
        }
      }
      for (int yl = 0; yl < 4; yl++)
        for (int xl = 0; xl < 4; xl++) {
          const int64_t y = by * 4 + yl, x = bx * 4 + xl;
          const uint8_t *p = &rgb[(y * width + x) * 3];
          int64_t r = p[0], g = p[1], b = p[2], a;
          const int64_t proj = dx * (2 * xl - 3) + dy * (2 * yl - 3);
          const int64_t den = pmax - pmin > 1 ? pmax - pmin : 1;
          a = aroll < 75 ? 255 : ((proj - pmin) * 255) / den;
          if (flat) {
            r = (int64_t)(fcol & 0xFF);
            g = (int64_t)((fcol >> 8) & 0xFF);
            b = (int64_t)((fcol >> 16) & 0xFF);
            a = ((fcol >> 24) % 4) == 0 ? (int64_t)((fcol >> 32) & 0xFF) : 255;
          }
          if (edge) {
            const int64_t edir = (int64_t)((he >> 8) % 16);
            const int64_t eoff = (int64_t)((he >> 16) % 5) - 2;
            const int side =
                kDirs[edir][0] * (2 * xl - 3) + kDirs[edir][1] * (2 * yl - 3) > eoff * 32;
            const uint64_t sel = side ? c1 : c0;
            r = (int64_t)(sel & 0xFF);
            g = (int64_t)((sel >> 8) & 0xFF);
            b = (int64_t)((sel >> 16) & 0xFF);
          }
          out[y * width + x] =
              (uint32_t)r | ((uint32_t)g << 8) | ((uint32_t)b << 16) | ((uint32_t)a << 24);
        }
    }
  free(rgb);
  return 0;
}
