// BC7 (BPTC) block encoder for gfx950.
//
// Reproduces FasTC's scalar BC7 search (BPTCEncoder/src/Compressor.cpp,
// RGBAEndpoints.{h,cpp}) bit for bit at quality 0, and its simulated
// annealing (Compressor.cpp:538-630) with a counter-based random stream at
// quality > 0.
//
// Work mapping: one wavefront (64 lanes) owns four 4x4 blocks.
//   phase 1  the four blocks' 64 pixels are loaded (one pixel per lane, rows of
//            16 consecutive pixels, so each quarter-wave reads 64 contiguous
//            bytes) into LDS.
//   phase 2  BoxSelection (Compressor.cpp:1670-1750) block by block: lane i
//            scores two-subset shape i and three-subset shape i; a wave
//            argmin with lowest-index ties picks the shapes.
//   phase 3  lane (block = lane/16, slot = lane%16) runs one subset fit of
//            CompressClusters (Compressor.cpp:1752-1817): CompressCluster,
//            k-means, least squares, grid clamp and the annealing loop.
//   phase 4  one lane per block sums the subset errors per mode in the
//            reference's mode order, keeps the first strict minimum and
//            writes the 128-bit block (CompressionMode::Pack, :1096-1298).
//
// Uniform error metric only: all errors are then exact integers (< 2^24), so
// they are kept in uint32 and equal the reference's float/double sums.
#include <float.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.h"

namespace fh {
namespace bc7 {

enum { PB_SHARED = 0, PB_NOTSHARED = 1, PB_NONE = 2 };

struct ModeAttr {
  int8_t part_bits, subsets, idx_bits, alpha_idx_bits, color_prec, alpha_prec, has_rot,
      has_idx_mode, pbit_type;
};

// Compressor.cpp:170-203
__constant__ ModeAttr kModes[8] = {
    {4, 3, 3, 0, 4, 0, 0, 0, PB_NOTSHARED}, {6, 2, 3, 0, 6, 0, 0, 0, PB_SHARED},
    {6, 3, 2, 0, 5, 0, 0, 0, PB_NONE},      {6, 2, 2, 0, 7, 0, 0, 0, PB_NOTSHARED},
    {0, 1, 2, 3, 5, 6, 1, 1, PB_NONE},      {0, 1, 2, 2, 7, 8, 1, 0, PB_NONE},
    {0, 1, 4, 0, 7, 7, 0, 0, PB_NOTSHARED}, {6, 2, 2, 0, 5, 5, 0, 0, PB_NOTSHARED}};

// BC7 partition tables (same values as BPTCEncoder/include/FasTC/Shapes.h:24-53).
__constant__ uint16_t kShape2[64] = {
    0xcccc, 0x8888, 0xeeee, 0xecc8, 0xc880, 0xfeec, 0xfec8, 0xec80, 0xc800, 0xffec, 0xfe80,
    0xe800, 0xffe8, 0xff00, 0xfff0, 0xf000, 0xf710, 0x008e, 0x7100, 0x08ce, 0x008c, 0x7310,
    0x3100, 0x8cce, 0x088c, 0x3110, 0x6666, 0x366c, 0x17e8, 0x0ff0, 0x718e, 0x399c, 0xaaaa,
    0xf0f0, 0x5a5a, 0x33cc, 0x3c3c, 0x55aa, 0x9696, 0xa55a, 0x73ce, 0x13c8, 0x324c, 0x3bdc,
    0x6996, 0xc33c, 0x9966, 0x0660, 0x0272, 0x04e4, 0x4e40, 0x2720, 0xc936, 0x936c, 0x39c6,
    0x639c, 0x9336, 0x9cc6, 0x817e, 0xe718, 0xccf0, 0x0fcc, 0x7744, 0xee22};
__constant__ uint16_t kShape3[64][2] = {
    {0xfecc, 0xf600}, {0xffc8, 0x7300}, {0xff90, 0x3310}, {0xecce, 0x00ce}, {0xff00, 0xcc00},
    {0xcccc, 0xcc00}, {0xffcc, 0x00cc}, {0xffcc, 0x3300}, {0xff00, 0xf000}, {0xfff0, 0xf000},
    {0xfff0, 0xff00}, {0xcccc, 0x8888}, {0xeeee, 0x8888}, {0xeeee, 0xcccc}, {0xffec, 0xec80},
    {0x739c, 0x7310}, {0xfec8, 0xc800}, {0x39ce, 0x3100}, {0xfff0, 0xccc0}, {0xfccc, 0x0ccc},
    {0xeeee, 0xee00}, {0xff88, 0x7700}, {0xeec0, 0xcc00}, {0x7730, 0x3300}, {0x0cee, 0x00cc},
    {0xffcc, 0xfc88}, {0x6ff6, 0x0660}, {0xff60, 0x6600}, {0xcbbc, 0xc88c}, {0xf966, 0xf900},
    {0xceec, 0x0cc0}, {0xff10, 0x7310}, {0xff80, 0xec80}, {0xccce, 0x08ce}, {0xeccc, 0xec80},
    {0x6666, 0x4444}, {0x0ff0, 0x0f00}, {0x6db6, 0x4924}, {0x6bd6, 0x4294}, {0xcf3c, 0x0c30},
    {0xc3fc, 0x03c0}, {0xffaa, 0xff00}, {0xff00, 0x5500}, {0xfcfc, 0xcccc}, {0xcccc, 0x0c0c},
    {0xf6f6, 0x6666}, {0xaffa, 0x0ff0}, {0xfff0, 0x5550}, {0xfaaa, 0xf000}, {0xeeee, 0x0e0e},
    {0xf8f8, 0x8888}, {0xfff0, 0x9990}, {0xeeee, 0xe00e}, {0x8ff8, 0x8888}, {0xf666, 0xf000},
    {0xff00, 0x9900}, {0xff66, 0xff00}, {0xcccc, 0xc00c}, {0xcffc, 0xcccc}, {0xf000, 0x9000},
    {0x8888, 0x0808}, {0xfefe, 0xeeee}, {0xfffa, 0xfff0}, {0x7bde, 0x7310}};
// Anchor (fix-up) pixels (BPTCEncoder/src/AnchorTables.h:22-51).
__constant__ uint8_t kAnchor2[64] = {
    15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 2,  8,  2,  2, 8,
    8,  15, 2,  8,  2,  2,  8,  8,  2,  2,  15, 15, 6,  8,  2,  8,  15, 15, 2,  8,  2, 2,
    2,  15, 15, 6,  6,  2,  6,  8,  15, 15, 2,  2,  15, 15, 15, 15, 15, 2,  2,  15};
__constant__ uint8_t kAnchor3[2][64] = {
    {3,  3, 15, 15, 8, 3,  15, 15, 8,  8,  6,  6,  6,  5,  3,  3,  3,  3,  8, 15, 3,  3,
     6,  10, 5, 8,  8, 6,  8,  5,  15, 15, 8,  15, 3,  5,  6,  10, 8,  15, 15, 3, 15, 5,
     15, 15, 15, 15, 3, 15, 5,  5,  5,  8,  5,  10, 5,  10, 8,  13, 15, 12, 3,  3},
    {15, 8,  8,  3,  15, 15, 3,  8,  15, 15, 15, 15, 15, 15, 15, 8,  15, 8,  15, 3,  15, 8,
     15, 8,  3,  15, 6,  10, 15, 15, 10, 8,  15, 3,  15, 10, 10, 8,  9,  10, 6,  15, 8,  15,
     3,  6,  6,  8,  15, 3,  15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 3,  15, 15, 8}};
// Weight of the second endpoint for 1..4-bit indices (Compressor.cpp:159-168);
// the first endpoint's weight is 64 minus it.
__constant__ uint8_t kW[4][16] = {
    {0, 31, 64},
    {0, 21, 43, 64},
    {0, 9, 18, 27, 37, 46, 55, 64},
    {0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64}};
// Solid-block alpha-index watermark (Compressor.cpp:135-138).
__constant__ uint32_t kWM[9] = {0x32b92180, 0x32ba3080, 0x31103200, 0x28103c80, 0x32bb3080,
                                0x25903600, 0x3530b900, 0x3b32b180, 0x34b5b98};
__constant__ uint8_t kMode5Tab[256][2];
// CompressClusters' mode order (Compressor.cpp:1757).
__constant__ int kOrder[8] = {0, 2, 1, 3, 7, 4, 5, 6};
__constant__ uint8_t kDxt6Tab[256][2][3];

__device__ __forceinline__ uint32_t chan(uint32_t p, int k) { return (p >> (8 * k)) & 0xFFu; }

__device__ __forceinline__ uint32_t subset_mask(int shape, int ns, int s) {
  if (ns == 2) {
    const uint32_t m = kShape2[shape];
    return s ? m : (~m & 0xFFFFu);
  }
  if (ns == 3) {
    const uint32_t m0 = kShape3[shape][0], m1 = kShape3[shape][1];
    if (s == 0) return ~m0 & 0xFFFFu;
    if (s == 1) return m0 & ~m1 & 0xFFFFu;
    return m0 & m1;
  }
  return 0xFFFFu;
}

__device__ __forceinline__ int anchor_of(int s, int shape, int ns) {
  if (s == 1) return ns == 2 ? kAnchor2[shape] : kAnchor3[0][shape];
  if (s == 2) return kAnchor3[1][shape];
  return 0;
}

// Point seen by the RGB fit.  ptmode 0: the pixel.  ptmode 1..4: rotation
// r = ptmode-1 of the alpha variant: channel r-1 swapped with A, then A = 255
// (Compressor.cpp:668-692).
__device__ __forceinline__ uint32_t point_of(uint32_t pix, int ptmode) {
  if (ptmode == 0) return pix;
  const uint32_t r = (uint32_t)ptmode - 1;
  uint32_t p = pix | 0xFF000000u;
  if (r >= 1) {
    const uint32_t sh = 8 * (r - 1);
    p = (p & ~(0xFFu << sh)) | ((pix >> 24) << sh);
  }
  return p;
}

// QuantizeChannel (RGBAEndpoints.cpp:126-165).
__device__ __forceinline__ uint32_t quant_channel(uint32_t val, uint32_t mask, int pbit) {
  if (mask == 0xFFu) return val;
  if (mask == 0u) return 0xFFu;
  uint32_t prec = __popc(mask);
  const uint32_t step = 1u << (8 - prec);
  uint32_t lval = val & mask;
  uint32_t hval = lval + step;
  if (pbit >= 0) {
    prec++;
    lval |= (uint32_t)pbit << (8 - prec);
    hval |= (uint32_t)pbit << (8 - prec);
  }
  if (lval > val) {
    lval -= step;
    hval -= step;
  }
  lval |= lval >> prec;
  hval |= hval >> prec;
  const uint32_t l8 = lval & 0xFFu, h8 = hval & 0xFFu;
  const uint32_t dl = val > l8 ? val - l8 : l8 - val;
  const uint32_t dh = val > h8 ? val - h8 : h8 - val;
  return dl < dh ? l8 : h8;
}

// ToPixel on integer-valued endpoints (uint32(c + 0.5) is then c).
__device__ __forceinline__ uint32_t to_pixel_int(uint32_t p, uint32_t mask, int pbit) {
  uint32_t r = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) r |= quant_channel(chan(p, k), chan(mask, k), pbit) << (8 * k);
  return r;
}

// uint32(c + 0.5) with the add in double (RGBAEndpoints.cpp:169-172), exactly,
// for c in [0, 255].
__device__ __forceinline__ uint32_t round_half_up(float c) {
  const float f = floorf(c);
  return (uint32_t)f + ((c - f) >= 0.5f ? 1u : 0u);
}

__device__ __forceinline__ uint32_t to_pixel_f(const float p[4], uint32_t mask, int pbit) {
  uint32_t r = 0;
#pragma unroll
  for (int k = 0; k < 4; k++)
    r |= quant_channel(round_half_up(p[k]) & 0xFFu, chan(mask, k), pbit) << (8 * k);
  return r;
}

__device__ __forceinline__ float dot4(const float a[4], const float b[4]) {
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 4; k++) s += a[k] * b[k];
  return s;
}

// VectorBase::Length: sqrt in long double rounded to float equals the
// correctly rounded single-precision sqrt (the double rounding is innocuous
// since 64 >= 2*24+2).
__device__ __forceinline__ float len4(const float a[4]) { return sqrtf(dot4(a, a)); }

__device__ __forceinline__ void normalize4(float a[4]) {
  const float l = len4(a);
#pragma unroll
  for (int k = 0; k < 4; k++) a[k] /= l;
}

// QuantizedError with endpoints already quantised to qp1/qp2
// (RGBAEndpoints.cpp:190-310, uniform metric).  Indices are written at pixel
// positions (4 bits each) for the pixels in smask.
__device__ uint32_t qerr(const uint32_t px[16], uint32_t smask, int ptmode, uint32_t qp1,
                         uint32_t qp2, int nb, uint64_t &idx) {
  int a[4], b[4], dir[4];
  int lsq = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    a[k] = (int)chan(qp1, k);
    b[k] = (int)chan(qp2, k);
    dir[k] = b[k] - a[k];
    lsq += dir[k] * dir[k];
  }
  idx = 0;
  uint32_t total = 0;
  if (lsq == 0) {
#pragma unroll
    for (int i = 0; i < 16; i++) {
      if (!((smask >> i) & 1)) continue;
      uint32_t e = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int d = (int)chan(px[i], k) - a[k];
        e += (uint32_t)(d * d);
      }
      total += e;
    }
    return total;
  }
  const int prec = nb == 4 ? 2 : (nb == 8 ? 3 : 4);
  const float flsq = (float)lsq;
  const float kf = (float)(nb - 1);
#pragma unroll
  for (int i = 0; i < 16; i++) {
    if (!((smask >> i) & 1)) continue;
    const uint32_t pt = point_of(px[i], ptmode);
    int d = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) d += ((int)chan(pt, k) - a[k]) * dir[k];
    const float pct = (float)d / flsq;
    const float t = pct * kf;
    int j1 = (int)floorf(t);
    int j2 = (int)ceilf(t);
    j1 = j1 < 0 ? 0 : j1;
    j1 = j1 > nb - 1 ? nb - 1 : j1;
    j2 = j2 > nb - 1 ? nb - 1 : j2;
    const uint32_t pix = px[i];
    uint32_t e1 = 0;
    {
      const int w1 = kW[prec - 1][j1], w0 = 64 - w1;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int ip = (a[k] * w0 + b[k] * w1 + 32) >> 6;
        const int dd = (int)chan(pix, k) - ip;
        e1 += (uint32_t)(dd * dd);
      }
    }
    int best = j1;
    uint32_t me = e1;
    if (j2 > j1) {
      const int w1 = kW[prec - 1][j2], w0 = 64 - w1;
      uint32_t e2 = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int ip = (a[k] * w0 + b[k] * w1 + 32) >> 6;
        const int dd = (int)chan(pix, k) - ip;
        e2 += (uint32_t)(dd * dd);
      }
      if (e2 < e1) {
        best = j2;
        me = e2;
      }
    }
    total += me;
    idx |= (uint64_t)best << (4 * i);
  }
  return total;
}

// Per-lane description of one subset fit.
struct Fit {
  int mode;
  int rot;        // rotation (modes 4/5), 0 otherwise
  int idx_mode;   // index mode (mode 4), 0 otherwise
  int nbits_idx;  // GetNumberOfBitsPerIndex() for this fit
  uint32_t qmask;
  int pbit_type;
  int color_prec, alpha_prec;
  int npbc;
};

__device__ __forceinline__ int pbc_bit(int pbit_type, int combo, int which) {
  // GetPBitCombo (CompressionMode.h:243-250): kPBits = {0,0},{0,1},{1,0},{1,1}
  if (pbit_type == PB_SHARED) return combo ? 1 : 0;
  if (pbit_type == PB_NOTSHARED) return which == 0 ? ((combo % 4) >> 1) : ((combo % 4) & 1);
  return 0;
}

// GetQuantizationMask (CompressionMode.h:212-232).
__device__ __forceinline__ uint32_t qmask_of(int color_prec, int alpha_prec) {
  const int32_t seed = (int32_t)0x80000000u;
  const uint32_t cbits = (uint32_t)color_prec - 1;
  uint32_t m = ((uint32_t)(seed >> (24 + cbits)) & 0xFFu) |
               ((uint32_t)(seed >> (16 + cbits)) & 0xFF00u) |
               ((uint32_t)(seed >> (8 + cbits)) & 0xFF0000u);
  if (alpha_prec > 0) m |= (uint32_t)(seed >> (alpha_prec - 1)) & 0xFF000000u;
  return m;
}

// CompressSingleColor (Compressor.cpp:252-353) through the precomputed
// per-value endpoint table.  Returns the per-pixel error.
__device__ uint32_t single_color(const Fit &f, const uint32_t *sctab, uint32_t pixel,
                                 uint32_t &p1, uint32_t &p2, int &pbc) {
  uint32_t best = 0xFFFFFFFFu;
  const int bpi = f.nbits_idx - 1;
  for (int pbi = 0; pbi < f.npbc; pbi++) {
    uint32_t err = 0, q1 = 0, q2 = 0;
#pragma unroll
    for (int ci = 0; ci < 4; ci++) {
      const uint32_t val = chan(pixel, ci);
      const int nbits = ci == 3 ? f.alpha_prec : f.color_prec;
      uint32_t lo, hi, dist;
      if (nbits == 0) {
        lo = hi = 0xFFu;
        dist = 0xFFu - val;
      } else {
        int cfg = 0;
        if (f.pbit_type != PB_NONE)
          cfg = 1 + 2 * pbc_bit(f.pbit_type, pbi, 0) + pbc_bit(f.pbit_type, pbi, 1);
        const uint32_t e = sctab[(((nbits - 4) * 5 + cfg) * 4 + bpi) * 256 + val];
        lo = e & 0xFFu;
        hi = (e >> 8) & 0xFFu;
        dist = (e >> 16) & 0xFFu;
      }
      err += dist * dist;
      q1 |= lo << (8 * ci);
      q2 |= hi << (8 * ci);
    }
    if (err < best) {
      best = err;
      pbc = pbi;
      p1 = q1;
      p2 = q2;
    }
  }
  return best;
}

// MatrixSquare<float,4>::PowerMethod without the eigenvalue
// (Base/include/FasTC/MatrixSquare.h:44-105).
__device__ void power_method(const float m[16], float b[4]) {
  const float norm = 1.0f / sqrtf(4.0f);
#pragma unroll
  for (int i = 0; i < 4; i++) b[i] = norm;
  bool bad = false, fixed = false;
  int iters = 0;
  while (!fixed && ++iters < 5) {
    float nb[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      nb[j] = 0.0f;
#pragma unroll
      for (int i = 0; i < 4; i++) nb[j] += m[j * 4 + i] * b[i];
    }
    const float len = len4(nb);
    if ((double)len < 1e-10) {
      if (bad) return;
      b[0] = 1.0f;
      b[1] = 1.0f;
      normalize4(b);
      bad = true;
      continue;
    }
    normalize4(nb);
    if (fabs((double)(1.0f - dot4(b, nb))) < 1e-8) fixed = true;
#pragma unroll
    for (int i = 0; i < 4; i++) b[i] = nb[i];
  }
}

__device__ __forceinline__ void unpack4f(uint32_t p, float v[4]) {
#pragma unroll
  for (int k = 0; k < 4; k++) v[k] = (float)chan(p, k);
}

// GetPrincipalAxis with no eigenvalues (RGBAEndpoints.cpp:327-428),
// including the collinearity test's reads of default (-1,-1,-1,-1) entries
// past the last unique point (:364-371).
__device__ void principal_axis(const uint32_t px[16], uint32_t smask, int ptmode, int n,
                               const float avg[4], float axis[4]) {
  // unique points in first-occurrence order; test u2.. as they appear
  int nu = 0;
  float u0[4], dir[4];
  bool collinear = true;
  uint32_t u0p = 0;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    if (!((smask >> i) & 1)) continue;
    const uint32_t pi = point_of(px[i], ptmode);
    bool seen = false;
#pragma unroll
    for (int j = 0; j < i; j++)
      if (((smask >> j) & 1) && point_of(px[j], ptmode) == pi) seen = true;
    if (seen) continue;
    if (nu == 0) {
      u0p = pi;
      unpack4f(pi, u0);
    } else if (nu == 1) {
      float u1[4];
      unpack4f(pi, u1);
#pragma unroll
      for (int k = 0; k < 4; k++) dir[k] = u1[k] - u0[k];
      normalize4(dir);
    } else if (collinear) {
      float v[4];
      unpack4f(pi, v);
#pragma unroll
      for (int k = 0; k < 4; k++) v[k] -= u0[k];
      if (fabs(fabs((double)dot4(v, dir)) - (double)len4(v)) > 1e-7) collinear = false;
    }
    nu++;
  }
  (void)u0p;
  if (nu == 1) {
#pragma unroll
    for (int k = 0; k < 4; k++) axis[k] = 0.0f;
    return;
  }
  // the loop runs i = 2 .. n-1 over upts[]; entries past nu are (-1,-1,-1,-1)
  if (collinear && n > (nu > 2 ? nu : 2)) {
    float v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = -1.0f - u0[k];
    if (fabs(fabs((double)dot4(v, dir)) - (double)len4(v)) > 1e-7) collinear = false;
  }
  if (collinear) {
#pragma unroll
    for (int k = 0; k < 4; k++) axis[k] = dir[k];
    return;
  }
  float cov[16];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j <= i; j++) {
      float sum = 0.0f;
#pragma unroll
      for (int p = 0; p < 16; p++) {
        if (!((smask >> p) & 1)) continue;
        const uint32_t pt = point_of(px[p], ptmode);
        const float ti = (float)chan(pt, i) - avg[i];
        const float tj = (float)chan(pt, j) - avg[j];
        sum += ti * tj;
      }
      cov[i * 4 + j] = sum / 3.0f;
      cov[j * 4 + i] = cov[i * 4 + j];
    }
  power_method(cov, axis);
}

// Counter-based annealing stream (the oracle defines the same function).
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
__device__ __forceinline__ uint32_t fit_key(uint64_t seed, uint64_t block_id, uint32_t tag) {
  uint32_t k = mix32((uint32_t)seed ^ 0x2545F491u);
  k = mix32(k ^ (uint32_t)(seed >> 32));
  k = mix32(k ^ (uint32_t)block_id);
  k = mix32(k ^ (uint32_t)(block_id >> 32));
  return mix32(k ^ tag);
}

struct Rng {
  uint32_t key, ctr;
  __device__ __forceinline__ uint32_t next16() { return mix32(key + (ctr++) * 0x9E3779B9u) >> 16; }
  // frand (Compressor.cpp:506-519)
  __device__ __forceinline__ float frand() {
    const uint32_t r = next16();
    return __uint_as_float((127u << 23) | ((r << 8) | (r >> 7))) - 1.0f;
  }
};

__device__ __forceinline__ uint32_t step_channel(uint32_t v, int s, bool minus) {
  int x = (int)v + (minus ? -s : s);
  x = x < 0 ? 0 : x;
  x = x > 255 ? 255 : x;
  return (uint32_t)x;
}

// PickBestNeighboringEndpoints (Compressor.cpp:426-498) on integer endpoints.
__device__ __forceinline__ void pick_neighbor(const Fit &f, Rng &rng, uint32_t p1, uint32_t p2,
                                              int cur_pbc, uint32_t &np1, uint32_t &np2,
                                              int &npbc, uint32_t vp1, uint32_t vp2, int vpbc) {
  int step[4];
  step[0] = step[1] = step[2] = 1 << (8 - f.color_prec);
  step[3] = 1 << (8 - f.alpha_prec);
  if (f.mode < 4) step[(f.rot + 3) % 4] = 0;
  const bool has_pb = f.pbit_type != PB_NONE;
  if (has_pb) npbc = f.pbit_type == PB_SHARED ? (cur_pbc + 1) % 2 : 3 - cur_pbc;
  bool visited = true;
  int guard = -1;
  while (visited && ++guard < 16) {
#pragma unroll
    for (int pt = 0; pt < 2; pt++) {
      const uint32_t p = pt ? p1 : p2;
      const uint32_t dirv = rng.next16() % 16;
      uint32_t np = 0;
      if (has_pb) {
        const int oldp = pbc_bit(f.pbit_type, cur_pbc, pt);
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const bool bit = (dirv >> k) & 1;
          uint32_t c = chan(p, k);
          if (bit && oldp == 0) c = step_channel(c, step[k], true);
          else if (!bit && oldp == 1) c = step_channel(c, step[k], false);
          np |= c << (8 * k);
        }
      } else {
#pragma unroll
        for (int k = 0; k < 4; k++)
          np |= step_channel(chan(p, k), step[k], (dirv >> k) & 1) << (8 * k);
      }
      if (pt) np1 = np;
      else np2 = np;
    }
    visited = np1 == vp1 && np2 == vp2 && npbc == vpbc;
  }
}

// OptimizeEndpointsForCluster (Compressor.cpp:538-630).  p1/p2 come in as
// floats (grid-clamped) and leave as the best integer endpoints.
__device__ uint32_t optimize(const Fit &f, uint32_t q, Rng &rng, const uint32_t px[16],
                             uint32_t smask, int ptmode, const float fp1[4], const float fp2[4],
                             uint32_t &bp1o, uint32_t &bp2o, uint64_t &best_idx, int &best_pbc) {
  const int nb = 1 << f.nbits_idx;
  // QuantizedError always receives a p-bit combo here (GetPBitCombo never
  // returns NULL), so modes without p-bits quantise with p-bit 0.
  uint32_t q1 = to_pixel_f(fp1, f.qmask, pbc_bit(f.pbit_type, best_pbc, 0));
  uint32_t q2 = to_pixel_f(fp2, f.qmask, pbc_bit(f.pbit_type, best_pbc, 1));
  uint32_t cur = qerr(px, smask, ptmode, q1, q2, nb, best_idx);
  int cur_pbc = best_pbc;
  uint32_t best = cur;
  uint32_t p1, p2;
  if (f.pbit_type != PB_NONE) {
    p1 = q1;
    p2 = q2;
  } else {
    p1 = to_pixel_f(fp1, f.qmask, -1);
    p2 = to_pixel_f(fp2, f.qmask, -1);
  }
  uint32_t bp1 = p1, bp2 = p2, vp1 = p1, vp2 = p2;
  int vpbc = cur_pbc;
  const int max_energy = (int)q;
  for (int energy = 0; best > 0 && energy < max_energy; energy++) {
    const float temp = (float)energy / (float)(max_energy - 1);
    uint32_t np1 = 0, np2 = 0;
    int npbc = 0;
    pick_neighbor(f, rng, p1, p2, cur_pbc, np1, np2, npbc, vp1, vp2, vpbc);
    const uint32_t e1 = to_pixel_int(np1, f.qmask, pbc_bit(f.pbit_type, npbc, 0));
    const uint32_t e2 = to_pixel_int(np2, f.qmask, pbc_bit(f.pbit_type, npbc, 1));
    uint64_t idx;
    const uint32_t err = qerr(px, smask, ptmode, e1, e2, nb, idx);
    // AcceptNewEndpointError (Compressor.cpp:524-536)
    bool accept = err < cur;
    if (!accept) {
      const double pr = exp(((double)0.1f * ((double)cur - (double)err)) / (double)temp);
      const double r = (double)rng.frand();
      accept = r < pr;
    }
    if (accept) {
      cur = err;
      p1 = np1;
      p2 = np2;
      cur_pbc = npbc;
    }
    if (err < best) {
      best_idx = idx;
      bp1 = np1;
      bp2 = np2;
      best_pbc = npbc;
      best = err;
      vp1 = np1;
      vp2 = np2;
      vpbc = npbc;
      energy = 0;
    }
  }
  bp1o = bp1;
  bp2o = bp2;
  return best;
}

// CompressCluster, RGB variant (Compressor.cpp:921-1094).  avg/mn/mx are the
// cluster statistics the reference uses (stale ones for the rotation copy).
// Returns the error; endpoints are integer-valued on return.
template <int NB>
__device__ uint32_t kmeans_lsq_fit(const Fit &f, uint32_t q, Rng &rng, const uint32_t *sctab,
                                   const uint32_t px[16], uint32_t smask, int ptmode, int n,
                                   const float avg[4], uint32_t &p1o, uint32_t &p2o,
                                   uint64_t &idx, int &pbc) {
  float axis[4];
  principal_axis(px, smask, ptmode, n, avg, axis);
  float mindp = FLT_MAX, maxdp = -FLT_MAX;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    if (!((smask >> i) & 1)) continue;
    float v[4];
    unpack4f(point_of(px[i], ptmode), v);
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] -= avg[k];
    const float dp = dot4(v, axis);
    if (dp < mindp) mindp = dp;
    if (dp > maxdp) maxdp = dp;
  }
  float p1[4], p2[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    p1[k] = avg[k] + axis[k] * mindp;
    p2[k] = avg[k] + axis[k] * maxdp;
    // ClampEndpoints (RGBAEndpoints.cpp:436-441)
    p1[k] = p1[k] < 0.0f ? 0.0f : (p1[k] > 255.0f ? 255.0f : p1[k]);
    p2[k] = p2[k] < 0.0f ? 0.0f : (p2[k] > 255.0f ? 255.0f : p2[k]);
  }
  float c[NB][4];
#pragma unroll
  for (int i = 0; i < NB; i++) {
    const float s = (float)i / (float)(NB - 1);
#pragma unroll
    for (int k = 0; k < 4; k++) c[i][k] = p1[k] * (1.0f - s) + p2[k] * s;
  }
  uint32_t bidx[16];
#pragma unroll
  for (int i = 0; i < 16; i++) bidx[i] = 0;
  uint32_t cnt[NB];
  bool fixed = false;
  int guard = 0;
  while (!fixed && guard++ < 4096) {
#pragma unroll
    for (int i = 0; i < 16; i++) {
      if (!((smask >> i) & 1)) continue;
      float v[4];
      unpack4f(point_of(px[i], ptmode), v);
      int minb = -1;
      float mind = FLT_MAX;
#pragma unroll
      for (int j = 0; j < NB; j++) {
        float d[4];
#pragma unroll
        for (int k = 0; k < 4; k++) d[k] = v[k] - c[j][k];
        const float ds = dot4(d, d);
        if (ds < mind) {
          mind = ds;
          minb = j;
        }
      }
      bidx[i] = (uint32_t)minb;
    }
    fixed = true;
#pragma unroll
    for (int j = 0; j < NB; j++) {
      uint32_t cn = 0;
      float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int i = 0; i < 16; i++) {
        if (((smask >> i) & 1) && bidx[i] == (uint32_t)j) {
          cn++;
          float v[4];
          unpack4f(point_of(px[i], ptmode), v);
#pragma unroll
          for (int k = 0; k < 4; k++) s[k] += v[k];
        }
      }
      if (cn != 0) {
#pragma unroll
        for (int k = 0; k < 4; k++) s[k] = s[k] / (float)cn;
      }
      if (s[0] != c[j][0] || s[1] != c[j][1] || s[2] != c[j][2] || s[3] != c[j][3]) fixed = false;
#pragma unroll
      for (int k = 0; k < 4; k++) c[j][k] = s[k];
      cnt[j] = cn;
    }
  }
  int nfilled = 0, last = 0;
#pragma unroll
  for (int j = 0; j < NB; j++)
    if (cnt[j] > 0) {
      nfilled++;
      last = j;
    }
  if (nfilled == 1) {
    float cl[4];
#pragma unroll
    for (int j = 0; j < NB; j++)
      if (j == last) {
#pragma unroll
        for (int k = 0; k < 4; k++) cl[k] = c[j][k];
      }
    const uint32_t e = single_color(f, sctab, to_pixel_f(cl, 0xFFFFFFFFu, -1), p1o, p2o, pbc);
    idx = 0;
#pragma unroll
    for (int i = 0; i < 16; i++)
      if ((smask >> i) & 1) idx |= (uint64_t)1 << (4 * i);
    return (uint32_t)n * e;
  }
  float asq = 0.0f, bsq = 0.0f, ab = 0.0f;
  float ax[4] = {0.0f, 0.0f, 0.0f, 0.0f}, bx[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int i = 0; i < NB; i++) {
    const float fbi = (float)(NB - 1 - i);
    const float fb = (float)(NB - 1);
    const float fi = (float)i;
    const float fn = (float)cnt[i];
    const float a = fbi / fb;
    const float b = fi / fb;
    asq += fn * a * a;
    bsq += fn * b * b;
    ab += fn * a * b;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      ax[k] += c[i][k] * a * fn;
      bx[k] += c[i][k] * b * fn;
    }
  }
  const float fdet = 1.0f / (asq * bsq - ab * ab);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    p1[k] = (ax[k] * bsq - bx[k] * ab) * fdet;
    p2[k] = (bx[k] * asq - ax[k] * ab) * fdet;
  }
  // ClampEndpointsToGrid (Compressor.cpp:212-250)
#pragma unroll
  for (int k = 0; k < 4; k++) {
    p1[k] = p1[k] < 0.0f ? 0.0f : (p1[k] > 255.0f ? 255.0f : p1[k]);
    p2[k] = p2[k] < 0.0f ? 0.0f : (p2[k] > 255.0f ? 255.0f : p2[k]);
  }
  {
    const bool has_pb = f.npbc > 1;
    float mind = FLT_MAX;
    float g1[4] = {-1.0f, -1.0f, -1.0f, -1.0f}, g2[4] = {-1.0f, -1.0f, -1.0f, -1.0f};
    for (int i = 0; i < f.npbc; i++) {
      const uint32_t q1 = to_pixel_f(p1, f.qmask, has_pb ? pbc_bit(f.pbit_type, i, 0) : -1);
      const uint32_t q2 = to_pixel_f(p2, f.qmask, has_pb ? pbc_bit(f.pbit_type, i, 1) : -1);
      float n1[4], n2[4], d1[4], d2[4];
      unpack4f(q1, n1);
      unpack4f(q2, n2);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        d1[k] = n1[k] - p1[k];
        d2[k] = n2[k] - p2[k];
      }
      const float dist = dot4(d1, d1) + dot4(d2, d2);
      if (dist < mind) {
        mind = dist;
#pragma unroll
        for (int k = 0; k < 4; k++) {
          g1[k] = n1[k];
          g2[k] = n2[k];
        }
        pbc = i;
      }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      p1[k] = g1[k];
      p2[k] = g2[k];
    }
  }
  return optimize(f, q, rng, px, smask, ptmode, p1, p2, p1o, p2o, idx, pbc);
}

__device__ uint32_t fit_rgb(const Fit &f, uint32_t q, Rng &rng, const uint32_t *sctab,
                            const uint32_t px[16], uint32_t smask, int ptmode, int n,
                            const float avg[4], bool all_same, uint32_t &p1, uint32_t &p2,
                            uint64_t &idx, int &pbc) {
  if (all_same) {
    // GetPoint(0): the first point of the subset
    const int first = __ffs(smask) - 1;
    uint32_t pt = 0;
#pragma unroll
    for (int i = 0; i < 16; i++)
      if (i == first) pt = point_of(px[i], ptmode);
    const uint32_t e = single_color(f, sctab, pt, p1, p2, pbc);
    idx = 0;
#pragma unroll
    for (int i = 0; i < 16; i++)
      if ((smask >> i) & 1) idx |= (uint64_t)1 << (4 * i);
    return (uint32_t)n * e;
  }
  if (f.nbits_idx == 2)
    return kmeans_lsq_fit<4>(f, q, rng, sctab, px, smask, ptmode, n, avg, p1, p2, idx, pbc);
  if (f.nbits_idx == 3)
    return kmeans_lsq_fit<8>(f, q, rng, sctab, px, smask, ptmode, n, avg, p1, p2, idx, pbc);
  return kmeans_lsq_fit<16>(f, q, rng, sctab, px, smask, ptmode, n, avg, p1, p2, idx, pbc);
}

// Result of one slot, kept in LDS for the pack phase.
struct FitRec {
  uint32_t err;   // 0xFFFFFFFF: slot unused
  uint32_t px1, px2;
  uint32_t cidx_lo, cidx_hi;  // color indices, 4 bits per pixel
  uint32_t aidx_lo, aidx_hi;  // alpha indices (rotation modes), 4 bits per pixel
  uint32_t meta;              // pbc | rot << 8 | idx_mode << 16
};

struct BlockSel {
  int cls;     // 0 solid, 1 transparent, 2 searched
  int layout;  // 0: modes 0/1/2/3/6/7 layout, 1: alpha layout (modes 4/5/6/7)
  uint32_t modes;
  int nshapes;
  int shape2, shape3;
};

// Slot -> fit.  Layout 0: 0-2 mode 0, 3-5 mode 2, 6-7 mode 1, 8-9 mode 3,
// 10-11 mode 7, 12 mode 6 (shape entry 0), 13 mode 6 (shape entry 1).
// Layout 1: 0-7 mode 4 (rot, idx mode), 8-11 mode 5 (rot), 12 mode 6,
// 13-14 mode 7.
__device__ __forceinline__ bool slot_fit(const BlockSel &s, int slot, uint32_t q, int &mode,
                                         int &entry, int &subset, int &rot, int &im) {
  rot = 0;
  im = 0;
  subset = 0;
  entry = 0;
  if (s.layout == 0) {
    if (slot < 6) {
      mode = slot < 3 ? 0 : 2;
      subset = slot % 3;
      entry = 1;
      if (s.nshapes < 2) return false;
      if (mode == 0 && s.shape3 >= 16) return false;
    } else if (slot < 12) {
      mode = slot < 8 ? 1 : (slot < 10 ? 3 : 7);
      subset = slot & 1;
      entry = 0;
    } else if (slot < 14) {
      mode = 6;
      entry = slot - 12;
      if (entry >= s.nshapes) return false;
      // Both shape entries fit the same 16-pixel cluster; without annealing
      // the second fit repeats the first exactly and cannot win (strict <).
      if (entry == 1 && q == 0) return false;
    } else {
      return false;
    }
  } else {
    if (slot < 8) {
      mode = 4;
      rot = slot >> 1;
      im = slot & 1;
    } else if (slot < 12) {
      mode = 5;
      rot = slot - 8;
    } else if (slot == 12) {
      mode = 6;
    } else if (slot < 15) {
      mode = 7;
      subset = slot - 13;
    } else {
      return false;
    }
  }
  return (s.modes >> mode) & 1;
}

// 128-bit LSB-first writer (BitStream::WriteBits, BitStream.h:65-91).
struct BitWriter {
  uint64_t lo = 0, hi = 0;
  int pos = 0;
  __device__ __forceinline__ void put(uint32_t v, int n) {
    if (n <= 0) return;
    const uint64_t vv = (uint64_t)(n >= 32 ? v : (v & ((1u << n) - 1u)));
    if (pos >= 64) {
      hi |= vv << (pos - 64);
    } else {
      lo |= vv << pos;
      if (pos + n > 64) hi |= vv >> (64 - pos);
    }
    pos += n;
  }
};

__device__ __forceinline__ uint32_t idx_at(uint32_t lo, uint32_t hi, int i) {
  return i < 8 ? (lo >> (4 * i)) & 0xFu : (hi >> (4 * (i - 8))) & 0xFu;
}

__device__ void write_block(uint8_t *dst, const BitWriter &w) {
  uint4 v;
  v.x = (uint32_t)w.lo;
  v.y = (uint32_t)(w.lo >> 32);
  v.z = (uint32_t)w.hi;
  v.w = (uint32_t)(w.hi >> 32);
  *reinterpret_cast<uint4 *>(dst) = v;
}

// CompressOptimalColorBC7 (Compressor.cpp:1424-1458).
__device__ void pack_solid(uint32_t pixel, uint32_t wm, BitWriter &w) {
  w.put(1u << 5, 6);
  w.put(0, 2);
#pragma unroll
  for (int ch = 0; ch < 3; ch++) {
    const uint32_t v = chan(pixel, ch);
    w.put(kMode5Tab[v][0], 7);
    w.put(kMode5Tab[v][1], 7);
  }
  w.put(pixel >> 24, 8);
  w.put(pixel >> 24, 8);
  w.put(0xaaaaaaabu, 31);
  w.put(kWM[wm], 31);
}

// CompressionMode::Pack (Compressor.cpp:1096-1298) from per-subset records.
// The p-bits are written unswapped after an anchor fix-up swap, as the
// reference does (:1149-1170 vs :1234-1241).
__device__ void pack_mode(int mode, int shape, const FitRec *rec, BitWriter &w) {
  const ModeAttr a = kModes[mode];
  const int ns = a.subsets;
  const int rot = (rec[0].meta >> 8) & 0xFF;
  const int im = (rec[0].meta >> 16) & 0xFF;
  w.put(1u << mode, mode + 1);
  w.put((uint32_t)shape, a.part_bits);
  if (a.has_rot) w.put((uint32_t)rot, 2);
  if (a.has_idx_mode) w.put((uint32_t)im, 1);
  uint32_t px1[3], px2[3];
  uint32_t cl = 0, ch = 0;  // color indices at pixel positions
  uint32_t al = rec[0].aidx_lo, ah = rec[0].aidx_hi;
  const int nib = (im == 0) ? a.idx_bits : a.alpha_idx_bits;
  const int nab = (im == 0) ? a.alpha_idx_bits : a.idx_bits;
  for (int s = 0; s < ns; s++) {
    px1[s] = rec[s].px1;
    px2[s] = rec[s].px2;
    cl |= rec[s].cidx_lo;
    ch |= rec[s].cidx_hi;
  }
  for (int s = 0; s < ns; s++) {
    const uint32_t smask = subset_mask(shape, ns, s);
    const int anchor = anchor_of(s, shape, ns);
    if (idx_at(cl, ch, anchor) >> (nib - 1)) {
      const uint32_t t = px1[s];
      px1[s] = px2[s];
      px2[s] = t;
      const uint32_t nv = (1u << nib) - 1;
      // invert this subset's color indices (all 16 entries in the reference,
      // but the others belong to other subsets and are never written)
      uint32_t nl = 0, nh = 0;
      for (int i = 0; i < 16; i++) {
        uint32_t v = idx_at(cl, ch, i);
        if ((smask >> i) & 1) v = nv - v;
        if (i < 8) nl |= v << (4 * i);
        else nh |= v << (4 * (i - 8));
      }
      cl = nl;
      ch = nh;
      if (a.has_rot) {
        const uint32_t nav = (1u << nab) - 1;
        uint32_t ml = 0, mh = 0;
        for (int i = 0; i < 16; i++) {
          const uint32_t v = nav - idx_at(al, ah, i);
          if (i < 8) ml |= v << (4 * i);
          else mh |= v << (4 * (i - 8));
        }
        al = ml;
        ah = mh;
      }
    }
    if (a.has_rot && (idx_at(al, ah, anchor) >> (nab - 1)) > 0) {
      const uint32_t a1 = px1[s] >> 24, a2 = px2[s] >> 24;
      px1[s] = (px1[s] & 0x00FFFFFFu) | (a2 << 24);
      px2[s] = (px2[s] & 0x00FFFFFFu) | (a1 << 24);
      const uint32_t nav = (1u << nab) - 1;
      uint32_t ml = 0, mh = 0;
      for (int i = 0; i < 16; i++) {
        const uint32_t v = nav - idx_at(al, ah, i);
        if (i < 8) ml |= v << (4 * i);
        else mh |= v << (4 * (i - 8));
      }
      al = ml;
      ah = mh;
    }
  }
  const int cp = a.color_prec, ap = a.alpha_prec;
  for (int c = 0; c < 3; c++)
    for (int s = 0; s < ns; s++) {
      w.put(chan(px1[s], c) >> (8 - cp), cp);
      w.put(chan(px2[s], c) >> (8 - cp), cp);
    }
  if (ap) {
    for (int s = 0; s < ns; s++) {
      w.put(chan(px1[s], 3) >> (8 - ap), ap);
      w.put(chan(px2[s], 3) >> (8 - ap), ap);
    }
  }
  if (a.pbit_type != PB_NONE) {
    for (int s = 0; s < ns; s++) {
      const int pbc = rec[s].meta & 0xFF;
      w.put((uint32_t)pbc_bit(a.pbit_type, pbc, 0), 1);
      if (a.pbit_type != PB_SHARED) w.put((uint32_t)pbc_bit(a.pbit_type, pbc, 1), 1);
    }
  }
  if (a.has_idx_mode && im == 1) {
    for (int i = 0; i < 16; i++) w.put(idx_at(al, ah, i), i == 0 ? 1 : 2);
    for (int i = 0; i < 16; i++) w.put(idx_at(cl, ch, i), i == 0 ? 2 : 3);
  } else {
    for (int i = 0; i < 16; i++) {
      int s = 0;
      for (int t = 1; t < ns; t++)
        if ((subset_mask(shape, ns, t) >> i) & 1) s = t;
      const int anchor = anchor_of(s, shape, ns);
      w.put(idx_at(cl, ch, i), i == anchor ? nib - 1 : nib);
    }
    if (a.has_rot)
      for (int i = 0; i < 16; i++) w.put(idx_at(al, ah, i), i == 0 ? nab - 1 : nab);
  }
}

__device__ __forceinline__ void set_fit(Fit &f, int mode, int rot, int im) {
  const ModeAttr a = kModes[mode];
  f.mode = mode;
  f.rot = rot;
  f.idx_mode = im;
  f.nbits_idx = im == 0 ? a.idx_bits : a.alpha_idx_bits;
  f.color_prec = a.color_prec;
  f.alpha_prec = a.alpha_prec;
  f.qmask = qmask_of(a.color_prec, a.alpha_prec);
  f.pbit_type = a.pbit_type;
  f.npbc = a.pbit_type == PB_SHARED ? 2 : (a.pbit_type == PB_NOTSHARED ? 4 : 1);
}

// The alpha/rotation variant of CompressCluster (Compressor.cpp:632-919) for a
// full 16-pixel cluster.  stats of the unrotated cluster feed the RGB fit.
__device__ uint32_t fit_alpha(const Fit &f, uint32_t q, Rng &rng, const uint32_t *sctab,
                              const uint32_t px[16], const float avg[4], uint32_t &out_px1,
                              uint32_t &out_px2, uint64_t &cidx, uint64_t &aidx) {
  const int ptmode = f.rot + 1;
  float avals[16];
  float amin = FLT_MAX, amax = -FLT_MAX;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const uint32_t src = f.rot == 0 ? 3 : (uint32_t)(f.rot - 1);
    avals[i] = (float)chan(px[i], (int)src);
    amin = avals[i] < amin ? avals[i] : amin;
    amax = amax < avals[i] ? avals[i] : amax;
  }
  uint32_t rp1, rp2;
  int dummy_pbc = 0;
  // the source cluster is not a single point here (solid blocks never reach
  // the search), so the stale AllSamePoint test is false
  const uint32_t rgb_err =
      fit_rgb(f, q, rng, sctab, px, 0xFFFFu, ptmode, 16, avg, false, rp1, rp2, cidx, dummy_pbc);

  const int abits = f.idx_mode == 0 ? kModes[f.mode].alpha_idx_bits : kModes[f.mode].idx_bits;
  const int nb = 1 << abits;
  float a1 = amin, a2 = amax;
  uint32_t alpha_err = 0;
  aidx = 0;
  if (a1 == a2) {
    const uint32_t a1be = (uint32_t)a1, a2be = (uint32_t)a2;
    if (f.mode == 5) {
      alpha_err = 0;
    } else {
      if (kDxt6Tab[a1be][0][0]) {
        a1 = (float)((kDxt6Tab[a1be][1][1] << 2) | (kDxt6Tab[a1be][0][1] >> 4));
        a2 = (float)((kDxt6Tab[a2be][1][2] << 2) | (kDxt6Tab[a2be][0][1] >> 4));
      } else {
        a1 = (float)((kDxt6Tab[a1be][0][1] << 2) | (kDxt6Tab[a1be][0][1] >> 4));
        a2 = (float)((kDxt6Tab[a2be][0][2] << 2) | (kDxt6Tab[a2be][0][1] >> 4));
      }
      const uint32_t fill = f.idx_mode == 1 ? 1u : 2u;
#pragma unroll
      for (int i = 0; i < 16; i++) aidx |= (uint64_t)fill << (4 * i);
      const int w1 = kW[abits - 1][fill], w0 = 64 - w1;
      const uint32_t ip = (((uint32_t)a1 * w0) + ((uint32_t)a2 * w1) + 32) >> 6 & 0xFFu;
      const uint32_t d = a1be > ip ? a1be - ip : ip - a1be;
      alpha_err = 16 * d * d;
    }
  } else {
    float vals[8];
#pragma unroll
    for (int i = 0; i < 8; i++) vals[i] = 0.0f;
#pragma unroll
    for (int i = 0; i < 8; i++)
      if (i < nb) vals[i] = amin + ((float)i / (float)(nb - 1)) * (amax - amin);
    uint32_t bk[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
      float md = 255.0f;
      bk[i] = 0;
#pragma unroll
      for (int j = 0; j < 8; j++) {
        if (j >= nb) continue;
        const float d = fabsf(avals[i] - vals[j]);
        if (d < md) {
          md = d;
          bk[i] = (uint32_t)j;
        }
      }
    }
    float npts[8];
    bool fixed = false;
    int guard = 0;
    while (!fixed && guard++ < 4096) {
      fixed = true;
#pragma unroll
      for (int j = 0; j < 8; j++) {
        float s = 0.0f, c = 0.0f;
#pragma unroll
        for (int i = 0; i < 16; i++)
          if (bk[i] == (uint32_t)j) {
            s += avals[i];
            c += 1.0f;
          }
        if (c > 0.0f) s /= c;
        npts[j] = c;
        if (j < nb) {
          fixed = fixed && (s == vals[j]);
        }
        vals[j] = s;
      }
#pragma unroll
      for (int i = 0; i < 16; i++) {
        float md = 255.0f;
#pragma unroll
        for (int j = 0; j < 8; j++) {
          if (j >= nb) continue;
          const float d = fabsf(avals[i] - vals[j]);
          if (d < md) {
            md = d;
            bk[i] = (uint32_t)j;
          }
        }
      }
    }
    float asq = 0.0f, bsq = 0.0f, ab = 0.0f, ax = 0.0f, bx = 0.0f;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      if (i >= nb) continue;
      const float fbi = (float)(nb - 1 - i);
      const float fb = (float)(nb - 1);
      const float fi = (float)i;
      const float a = fbi / fb;
      const float b = fi / fb;
      const float n = npts[i];
      const float x = vals[i];
      asq += n * a * a;
      bsq += n * b * b;
      ab += n * a * b;
      ax += x * a * n;
      bx += x * b * n;
    }
    const float fd = 1.0f / (asq * bsq - ab * ab);
    a1 = fd * (ax * bsq - bx * ab);
    a2 = fd * (bx * asq - ax * ab);
    {
      float t = 0.0f < a1 ? a1 : 0.0f;
      a1 = t < 255.0f ? t : 255.0f;
      t = 0.0f < a2 ? a2 : 0.0f;
      a2 = t < 255.0f ? t : 255.0f;
    }
    const uint32_t amask = (uint32_t)((int8_t)-0x7F >> (f.alpha_prec - 1)) & 0xFFu;
    const uint32_t a1b = quant_channel((uint32_t)a1 & 0xFFu, amask, -1);
    const uint32_t a2b = quant_channel((uint32_t)a2 & 0xFFu, amask, -1);
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const uint32_t val = (uint32_t)avals[i];
      uint32_t me = 0xFFFFFFFFu;
      uint32_t bb = 0;
#pragma unroll
      for (int j = 0; j < 8; j++) {
        if (j >= nb) continue;
        const int w1 = kW[abits - 1][j], w0 = 64 - w1;
        const uint32_t ip = ((a1b * w0) + (a2b * w1) + 32) >> 6 & 0xFFu;
        const uint32_t d = val > ip ? val - ip : ip - val;
        const uint32_t e = d * d;
        if (e < me) {
          me = e;
          bb = (uint32_t)j;
        }
      }
      alpha_err += me;
      aidx |= (uint64_t)bb << (4 * i);
    }
  }
  // Pack quantises (rgb of the RGB fit, fitted alpha) with rounding and no p-bit.
  float e1[4], e2[4];
  unpack4f(rp1, e1);
  unpack4f(rp2, e2);
  e1[3] = a1;
  e2[3] = a2;
  out_px1 = to_pixel_f(e1, f.qmask, -1);
  out_px2 = to_pixel_f(e2, f.qmask, -1);
  return rgb_err + alpha_err;
}

__global__ void __launch_bounds__(64) encode_kernel(BlockRange r, Bc7Params prm,
                                                    Bc7Scratch scr, const uint32_t *sctab) {
  __shared__ uint32_t s_px[4][16];
  __shared__ BlockSel s_sel[4];
  __shared__ FitRec s_rec[4][16];
  const int lane = threadIdx.x;
  const uint32_t base = r.first_block + blockIdx.x * 4;

  // phase 1: rows of 16 pixels over 4 consecutive blocks
  {
    const int row = lane >> 4, x = lane & 15;
    const int blk = x >> 2;
    const uint32_t b = base + blk;
    if (b < r.first_block + r.num_blocks) {
      const uint32_t bx = b % r.blocks_wide, by = b / r.blocks_wide;
      s_px[blk][row * 4 + (x & 3)] = r.img[(size_t)(by * 4 + row) * r.width + bx * 4 + (x & 3)];
    }
  }
  __syncthreads();

  // phase 2: classify and select shapes, one block at a time
  for (int blk = 0; blk < 4; blk++) {
    const uint32_t b = base + blk;
    if (b >= r.first_block + r.num_blocks) break;  // uniform
    uint32_t px[16];
#pragma unroll
    for (int i = 0; i < 16; i++) px[i] = s_px[blk][i];
    bool solid = true, transparent = true, opaque = true;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      solid = solid && px[i] == px[0];
      transparent = transparent && (px[i] >> 24) == 0;
      opaque = opaque && (px[i] >> 24) >= 250;
    }
    BlockSel sel;
    sel.cls = solid ? 0 : (transparent ? 1 : 2);
    sel.layout = 0;
    sel.modes = 0;
    sel.nshapes = 1;
    sel.shape2 = sel.shape3 = 0;
    if (sel.cls == 2) {
      // EstimateTwoClusterError / EstimateThreeClusterError per shape
      double err2 = 0.0, err3 = 0.0;
      for (int ns = 2; ns <= 3; ns++) {
        if (ns == 3 && !opaque) break;
        double err = 0.0;
        for (int s = 0; s < ns; s++) {
          const uint32_t m = subset_mask(lane, ns, s);
          uint32_t mn = 0xFFFFFFFFu, mx = 0;
          uint32_t mnc[4] = {255, 255, 255, 255}, mxc[4] = {0, 0, 0, 0};
#pragma unroll
          for (int i = 0; i < 16; i++) {
            if (!((m >> i) & 1)) continue;
#pragma unroll
            for (int k = 0; k < 4; k++) {
              const uint32_t c = chan(px[i], k);
              mnc[k] = c < mnc[k] ? c : mnc[k];
              mxc[k] = c > mxc[k] ? c : mxc[k];
            }
          }
          mn = mnc[0] | (mnc[1] << 8) | (mnc[2] << 16) | (mnc[3] << 24);
          mx = mxc[0] | (mxc[1] << 8) | (mxc[2] << 16) | (mxc[3] << 24);
          if (mn != mx) {
            uint64_t dummy;
            const uint32_t qe = qerr(px, m, 0, mn, mx, ns == 2 ? 8 : 4, dummy);
            err += 0.0001 + (double)qe;
          } else {
            err += 0.0;
          }
        }
        if (ns == 2) err2 = err;
        else err3 = err;
      }
      // wave argmin, lowest index on ties
      double v2 = err2;
      int i2 = lane;
      double v3 = err3;
      int i3 = lane;
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const double o2 = __shfl_xor(v2, off);
        const int oi2 = __shfl_xor(i2, off);
        if (o2 < v2 || (o2 == v2 && oi2 < i2)) {
          v2 = o2;
          i2 = oi2;
        }
        const double o3 = __shfl_xor(v3, off);
        const int oi3 = __shfl_xor(i3, off);
        if (o3 < v3 || (o3 == v3 && oi3 < i3)) {
          v3 = o3;
          i3 = oi3;
        }
      }
      sel.shape2 = i2;
      if (v2 < 1e-9) {
        sel.modes = (1u << 1) | (1u << 3) | (1u << 7);
        sel.nshapes = 1;
        sel.layout = 0;
      } else if (!opaque) {
        sel.modes = (1u << 4) | (1u << 5) | (1u << 6) | (1u << 7);
        sel.nshapes = 1;
        sel.layout = 1;
      } else {
        sel.modes = 0xFFu & ~((1u << 4) | (1u << 5));
        sel.nshapes = 2;
        sel.shape3 = i3;
        sel.layout = 0;
        if (v3 < 1e-9) sel.modes = (1u << 0) | (1u << 2);
      }
      sel.modes &= prm.block_modes;
    }
    if (lane == 0) s_sel[blk] = sel;
  }
  __syncthreads();

  // phase 3: one subset fit per lane
  {
    const int blk = lane >> 4, slot = lane & 15;
    const uint32_t b = base + blk;
    FitRec rec;
    rec.err = 0xFFFFFFFFu;
    rec.px1 = rec.px2 = 0;
    rec.cidx_lo = rec.cidx_hi = rec.aidx_lo = rec.aidx_hi = 0;
    rec.meta = 0;
    if (b < r.first_block + r.num_blocks) {
      const BlockSel sel = s_sel[blk];
      int mode, entry, subset, rot, im;
      if (sel.cls == 2 && slot_fit(sel, slot, prm.quality, mode, entry, subset, rot, im)) {
        uint32_t px[16];
#pragma unroll
        for (int i = 0; i < 16; i++) px[i] = s_px[blk][i];
        Fit f;
        set_fit(f, mode, rot, im);
        const int ns = kModes[mode].subsets;
        const int shape = ns == 3 ? sel.shape3 : sel.shape2;
        Rng rng;
        rng.key = fit_key(prm.seed, prm.block_id_base + b,
                          (uint32_t)mode | ((uint32_t)entry << 3) | ((uint32_t)subset << 5) |
                              ((uint32_t)rot << 7) | ((uint32_t)im << 9));
        rng.ctr = 0;
        uint64_t cidx = 0, aidx = 0;
        uint32_t err;
        int pbc = 0;
        if (kModes[mode].has_rot) {
          // stale statistics of the unrotated 16-pixel cluster
          float avg[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
          for (int i = 0; i < 16; i++)
#pragma unroll
            for (int k = 0; k < 4; k++) avg[k] += (float)chan(px[i], k);
#pragma unroll
          for (int k = 0; k < 4; k++) avg[k] = avg[k] / 16.0f;
          err = fit_alpha(f, prm.quality, rng, sctab, px, avg, rec.px1, rec.px2, cidx, aidx);
        } else {
          const uint32_t smask = subset_mask(shape, ns, subset);
          const int n = __popc(smask);
          float avg[4] = {0.0f, 0.0f, 0.0f, 0.0f};
          uint32_t mnc = 0xFFFFFFFFu, mxc = 0;
          bool same = true;
          uint32_t firstp = 0;
          bool have_first = false;
#pragma unroll
          for (int i = 0; i < 16; i++) {
            if (!((smask >> i) & 1)) continue;
            if (!have_first) {
              firstp = px[i];
              have_first = true;
            }
            same = same && px[i] == firstp;
#pragma unroll
            for (int k = 0; k < 4; k++) avg[k] += (float)chan(px[i], k);
          }
          (void)mnc;
          (void)mxc;
#pragma unroll
          for (int k = 0; k < 4; k++) avg[k] = avg[k] / (float)n;
          uint32_t p1, p2;
          err = fit_rgb(f, prm.quality, rng, sctab, px, smask, 0, n, avg, same, p1, p2, cidx, pbc);
          // Pack's ToPixel of the chosen endpoints (idempotent on grid values)
          if (f.pbit_type == PB_NONE) {
            rec.px1 = to_pixel_int(p1, f.qmask, -1);
            rec.px2 = to_pixel_int(p2, f.qmask, -1);
          } else {
            rec.px1 = to_pixel_int(p1, f.qmask, pbc_bit(f.pbit_type, pbc, 0));
            rec.px2 = to_pixel_int(p2, f.qmask, pbc_bit(f.pbit_type, pbc, 1));
          }
        }
        rec.err = err;
        rec.cidx_lo = (uint32_t)cidx;
        rec.cidx_hi = (uint32_t)(cidx >> 32);
        rec.aidx_lo = (uint32_t)aidx;
        rec.aidx_hi = (uint32_t)(aidx >> 32);
        rec.meta = (uint32_t)pbc | ((uint32_t)rot << 8) | ((uint32_t)im << 16);
      }
    }
    s_rec[blk][slot] = rec;
  }
  __syncthreads();

  // phase 4: pick the mode and pack, one lane per block
  if (lane < 4) {
    const int blk = lane;
    const uint32_t b = base + blk;
    if (b < r.first_block + r.num_blocks) {
      const BlockSel sel = s_sel[blk];
      BitWriter w;
      if (sel.cls == 0) {
        const uint32_t rel = b - r.first_block;
        const uint32_t wm = (scr.chunk_prefix[rel / 256] + scr.local_prefix[rel]) % 9;
        pack_solid(s_px[blk][0], wm, w);
      } else if (sel.cls == 1) {
        w.put(1u << 6, 7);  // WriteTransparentBlock (Compressor.cpp:1416-1421)
      } else {
        uint64_t best = 0xFFFFFFFFFFFFFFFFull;
        int bmode = -1, bfirst = 0, bshape = 0;
        for (int oi = 0; oi < 8; oi++) {
          const int mode = kOrder[oi];
          if (!((sel.modes >> mode) & 1)) continue;
          // candidates of this mode in the reference's evaluation order
          for (int c = 0; c < 8; c++) {
            int first = -1, nsub = 1, shape = 0;
            if (sel.layout == 0) {
              if (mode == 0 || mode == 2) {
                if (c > 0) break;
                first = mode == 0 ? 0 : 3;
                nsub = 3;
                shape = sel.shape3;
              } else if (mode == 1 || mode == 3 || mode == 7) {
                if (c > 0) break;
                first = mode == 1 ? 6 : (mode == 3 ? 8 : 10);
                nsub = 2;
                shape = sel.shape2;
              } else if (mode == 6) {
                if (c > 1) break;
                first = 12 + c;
                shape = c == 0 ? sel.shape2 : sel.shape3;
              } else {
                break;
              }
            } else {
              if (mode == 4) {
                first = c;
              } else if (mode == 5) {
                if (c > 3) break;
                first = 8 + c;
              } else if (mode == 6) {
                if (c > 0) break;
                first = 12;
                shape = sel.shape2;
              } else if (mode == 7) {
                if (c > 0) break;
                first = 13;
                nsub = 2;
                shape = sel.shape2;
              } else {
                break;
              }
            }
            uint64_t e = 0;
            bool ok = true;
            for (int s = 0; s < nsub; s++) {
              const uint32_t se = s_rec[blk][first + s].err;
              if (se == 0xFFFFFFFFu) ok = false;
              e += se;
            }
            if (!ok) continue;
            if (e < best) {
              best = e;
              bmode = mode;
              bfirst = first;
              bshape = shape;
            }
          }
        }
        if (bmode >= 0) pack_mode(bmode, bshape, &s_rec[blk][bfirst], w);
      }
      write_block(r.out + (size_t)b * 16, w);
    }
  }
}

// Solid-block flags and in-chunk exclusive prefix (chunks of 256 blocks).
__global__ void __launch_bounds__(256) solid_scan_kernel(BlockRange r, Bc7Scratch scr) {
  __shared__ uint32_t s_wave[4];
  const uint32_t t = threadIdx.x;
  const uint32_t rel = blockIdx.x * 256 + t;  // relative to the job's first block
  const uint32_t b = r.first_block + rel;
  bool solid = false;
  if (rel < r.num_blocks) {
    const uint32_t bx = b % r.blocks_wide, by = b / r.blocks_wide;
    const uint32_t *p = r.img + (size_t)(by * 4) * r.width + bx * 4;
    const uint32_t p0 = p[0];
    solid = true;
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const uint4 row = *reinterpret_cast<const uint4 *>(p + (size_t)y * r.width);
      solid = solid && row.x == p0 && row.y == p0 && row.z == p0 && row.w == p0;
    }
  }
  const uint64_t ballot = __ballot(solid);
  const uint32_t lanei = t & 63, wave = t >> 6;
  const uint32_t before = __popcll(ballot & ((1ull << lanei) - 1ull));
  if (lanei == 0) s_wave[wave] = __popcll(ballot);
  __syncthreads();
  uint32_t wbase = 0;
  for (uint32_t w = 0; w < wave; w++) wbase += s_wave[w];
  if (rel < r.num_blocks) scr.local_prefix[rel] = (uint8_t)(wbase + before);
  if (t == 0) scr.chunk_prefix[blockIdx.x] = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
}

// Exclusive scan of chunk counts (one workgroup), seeded with the watermark
// counter c: the k-th solid block of the job gets table entry (c + 1 + k) % 9
// (gWMVal = (gWMVal + 1) % 9 per solid block, Compressor.cpp:1457).  The
// counter is left at the last entry used, or unchanged if no block was solid.
__global__ void __launch_bounds__(1024) chunk_scan_kernel(Bc7Scratch scr) {
  __shared__ uint32_t s_part[1024];
  __shared__ uint32_t s_total;
  const uint32_t t = threadIdx.x;
  const uint32_t c0 = *scr.counter;
  const uint32_t g1 = (c0 + 1u) % 9u;
  if (t == 0) s_total = 0;
  __syncthreads();
  for (uint32_t base = 0; base < scr.num_chunks; base += 1024) {
    const uint32_t i = base + t;
    const uint32_t v = i < scr.num_chunks ? scr.chunk_prefix[i] : 0;
    s_part[t] = v;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
      const uint32_t add = t >= off ? s_part[t - off] : 0;
      __syncthreads();
      s_part[t] += add;
      __syncthreads();
    }
    const uint32_t excl = s_total + s_part[t] - v;
    if (i < scr.num_chunks) scr.chunk_prefix[i] = (g1 + excl % 9u) % 9u;
    __syncthreads();
    if (t == 1023) s_total += s_part[1023];
    __syncthreads();
  }
  if (t == 0 && s_total > 0) *scr.counter = (g1 + (s_total - 1u) % 9u) % 9u;
}

}  // namespace bc7

// Host launchers -------------------------------------------------------------

namespace {
int g_tables_uploaded[64];
}

hipError_t bc7_upload_tables(int device) {
  if (device < 0 || device >= 64) return hipErrorInvalidDevice;
  if (g_tables_uploaded[device]) return hipSuccess;
  uint8_t m5[256][2];
  uint8_t d6[256][2][3];
  build_bc7_mode5_table(m5);
  build_bc7_dxt6_table(d6);
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(bc7::kMode5Tab), m5, sizeof(m5));
  if (e != hipSuccess) return e;
  e = hipMemcpyToSymbol(HIP_SYMBOL(bc7::kDxt6Tab), d6, sizeof(d6));
  if (e != hipSuccess) return e;
  g_tables_uploaded[device] = 1;
  return hipSuccess;
}

hipError_t bc7_encode(const BlockRange &r, const Bc7Params &p, const Bc7Scratch &s,
                      hipStream_t stream) {
  if (r.num_blocks == 0) return hipSuccess;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = bc7_upload_tables(dev);
  if (e != hipSuccess) return e;
  const uint32_t *sct = bc7_single_color_table(dev);
  if (!sct) return hipErrorOutOfMemory;
  const uint32_t nchunks = (r.num_blocks + 255) / 256;
  hipLaunchKernelGGL(bc7::solid_scan_kernel, dim3(nchunks), dim3(256), 0, stream, r, s);
  hipLaunchKernelGGL(bc7::chunk_scan_kernel, dim3(1), dim3(1024), 0, stream, s);
  const uint32_t ngroups = (r.num_blocks + 3) / 4;
  hipLaunchKernelGGL(bc7::encode_kernel, dim3(ngroups), dim3(64), 0, stream, r, p, s, sct);
  return hipGetLastError();
}

}  // namespace fh
