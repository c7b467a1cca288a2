// BC7 (BPTC) block encoder for gfx950.
//
// Reproduces FasTC's scalar BC7 search (BPTCEncoder/src/Compressor.cpp,
// RGBAEndpoints.{h,cpp}) bit for bit at quality 0, and its simulated
// annealing (Compressor.cpp:538-630) with a counter-based random stream at
// quality > 0.
//
// The search runs as a pipeline of kernels, each shaped so that the lanes of
// a wavefront execute the same code:
//   select  one wave per four blocks: classify (solid / transparent / search)
//           and run BoxSelection (Compressor.cpp:1670-1750) with lane i scoring
//           two- and three-subset shape i; a wave argmin with lowest-index ties
//           picks the shapes.  Each block's candidates (one per mode and shape
//           entry of CompressClusters, :1752-1817, with the rotation modes split
//           per rotation and index mode) are appended to per-mode queues.
//   setup   persistent, one lane per candidate, queue by queue, so a wave runs
//           one mode: principal axis, k-means, least squares and grid clamp for
//           every subset (CompressCluster, :921-1094), and the alpha fit of the
//           rotation modes (:632-919).
//   anneal  persistent, one lane per candidate: each iteration advances every
//           still-annealing subset of the candidate by one step and scores all
//           16 pixels in one pass, each against its subset's endpoints.  Lanes
//           that finish refill from a wave-local window of the work list.
//   pack    one lane per block: sum the subset errors per mode in the
//           reference's mode order, keep the first strict minimum and write the
//           128-bit block (CompressionMode::Pack, :1096-1298); solid blocks get
//           the watermark entry from an exclusive scan of solid-block flags.
//
// Uniform error metric only: all errors are then exact integers (< 2^24), so
// they are kept in uint32 and equal the reference's float/double sums.
#include <float.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fasthip.h"
#include "bc7_block_decode.h"
#include "ranker_tile.h"
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

// The same attributes for compile-time specialisation of per-mode kernels.
constexpr ModeAttr kModeTab[8] = {
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
// 1 / n in double for the k-means means (index n = 1..16)
__constant__ double kInvCount[17] = {0.0,        1.0,        1.0 / 2,  1.0 / 3,  1.0 / 4,
                                     1.0 / 5,    1.0 / 6,    1.0 / 7,  1.0 / 8,  1.0 / 9,
                                     1.0 / 10,   1.0 / 11,   1.0 / 12, 1.0 / 13, 1.0 / 14,
                                     1.0 / 15,   1.0 / 16};

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
__host__ __device__ __forceinline__ uint32_t quant_channel(uint32_t val, uint32_t mask, int pbit) {
  if (mask == 0xFFu) return val;
  if (mask == 0u) return 0xFFu;
  uint32_t prec = (uint32_t)__builtin_popcount(mask);
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

// quant_channel of every byte for each channel precision (the number of
// bits of the channel's mask, which keeps the top bits: 0..8) and p-bit
// (-1, 0, 1), at [(bits * 3 + pbit + 1) * 256 + byte]; built on the host
// from quant_channel itself.
__constant__ uint8_t kQuant[9 * 3 * 256];

// quant_channel(val, the channel mask byte, pbit) for a top-aligned mask
__device__ __forceinline__ uint32_t quant_tab(uint32_t val, uint32_t mask_byte, int pbit) {
  return kQuant[(uint32_t)__builtin_popcount(mask_byte) * 768u + (uint32_t)(pbit + 1) * 256u + val];
}

// to_pixel_int through kQuant: a load per channel in place of the
// comparisons and shifts (the annealing step quantises both endpoints of
// every subset).
__device__ __forceinline__ uint32_t to_pixel_tab(uint32_t p, uint32_t mask, int pbit) {
  // the colour channels share one precision (qmask_of), alpha has its own
  const uint32_t pb = (uint32_t)(pbit + 1) * 256u;
  const uint32_t cb = (uint32_t)__builtin_popcount(mask & 0xFFu) * 768u + pb;
  const uint32_t ab = (uint32_t)__builtin_popcount(mask >> 24) * 768u + pb;
  return (uint32_t)kQuant[cb + chan(p, 0)] | ((uint32_t)kQuant[cb + chan(p, 1)] << 8) |
         ((uint32_t)kQuant[cb + chan(p, 2)] << 16) | ((uint32_t)kQuant[ab + chan(p, 3)] << 24);
}

// ToPixel on integer-valued endpoints (uint32(c + 0.5) is then c).
__device__ __forceinline__ uint32_t to_pixel_int(uint32_t p, uint32_t mask, int pbit) {
  uint32_t r = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) r |= quant_tab(chan(p, k), chan(mask, k), pbit) << (8 * k);
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
    r |= quant_tab(round_half_up(p[k]) & 0xFFu, chan(mask, k), pbit) << (8 * k);
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
// UNROLL: 16 when px is a register array (constant indices), 1 when it
// points to LDS (the loop stays rolled).
template <int UNROLL>
__device__ __forceinline__ uint32_t qerr(const uint32_t px[16], uint32_t smask, int ptmode,
                                         uint32_t qp1, uint32_t qp2, int nb, uint64_t &idx) {
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
#pragma unroll UNROLL
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

// ---------------------------------------------------------------------------
// The nonuniform error metric (m_ErrorMetric = eErrorMetric_Nonuniform,
// BPTCCompressor.h:73-77; weights kErrorMetrics, Compressor.cpp:205-208).
// Each channel's absolute difference is multiplied by its weight before
// squaring, and the four squares are summed in float in channel order from
// zero (VectorBase::Dot, VectorBase.h:96-101).  The errors are then floats
// whose values depend on the order of the additions, so every sum below runs
// in the reference's order: channels, then the pixels of a subset in pixel
// order.  A float error is kept as its bit pattern in the 32-bit error fields
// the uniform metric fills with integers: for errors >= 0 the bit patterns
// order as the values do, so comparisons and the "> 0" tests are shared.
// ---------------------------------------------------------------------------
struct Weights {
  float w[4];
};

__device__ __forceinline__ float werr(int d0, int d1, int d2, int d3, const Weights &m) {
  const float e0 = (float)abs(d0) * m.w[0];
  const float e1 = (float)abs(d1) * m.w[1];
  const float e2 = (float)abs(d2) * m.w[2];
  const float e3 = (float)abs(d3) * m.w[3];
  return ((e0 * e0 + e1 * e1) + e2 * e2) + e3 * e3;
}

// The weights a rotation mode sees (GetErrorMetric, CompressionMode.h:196-205):
// rotation r swaps channel r-1's weight with alpha's.
__device__ __forceinline__ Weights rotate_weights(const Weights &m, int rot) {
  Weights r = m;
  if (rot >= 1 && rot <= 3) {
    r.w[rot - 1] = m.w[3];
    r.w[3] = m.w[rot - 1];
  }
  return r;
}

__device__ __forceinline__ Weights params_weights(const Bc7Params &prm) {
  Weights m;
#pragma unroll
  for (int k = 0; k < 4; k++) m.w[k] = prm.metric[k];
  return m;
}

// QuantizedError under per-channel weights m (qerr's float counterpart):
// per pixel the weighted error of each of the two neighbouring indices, the
// second kept only when strictly better, summed in float in pixel order.
template <int UNROLL>
__device__ __forceinline__ float qerr_nu(const uint32_t px[16], uint32_t smask, int ptmode,
                                         uint32_t qp1, uint32_t qp2, int nb, const Weights &m,
                                         uint64_t &idx) {
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
  float total = 0.0f;
  if (lsq == 0) {
#pragma unroll
    for (int i = 0; i < 16; i++) {
      if (!((smask >> i) & 1)) continue;
      total += werr((int)chan(px[i], 0) - a[0], (int)chan(px[i], 1) - a[1],
                    (int)chan(px[i], 2) - a[2], (int)chan(px[i], 3) - a[3], m);
    }
    return total;
  }
  const int prec = nb == 4 ? 2 : (nb == 8 ? 3 : 4);
  const float flsq = (float)lsq;
  const float kf = (float)(nb - 1);
#pragma unroll UNROLL
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
    int dd[4];
    {
      const int w1 = kW[prec - 1][j1], w0 = 64 - w1;
#pragma unroll
      for (int k = 0; k < 4; k++) dd[k] = (int)chan(pix, k) - ((a[k] * w0 + b[k] * w1 + 32) >> 6);
    }
    const float e1 = werr(dd[0], dd[1], dd[2], dd[3], m);
    int best = j1;
    float me = e1;
    if (j2 > j1) {
      const int w1 = kW[prec - 1][j2], w0 = 64 - w1;
#pragma unroll
      for (int k = 0; k < 4; k++) dd[k] = (int)chan(pix, k) - ((a[k] * w0 + b[k] * w1 + 32) >> 6);
      const float e2 = werr(dd[0], dd[1], dd[2], dd[3], m);
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
// per-value endpoint table.  Returns the per-pixel error; under the
// nonuniform metric (NU) its float bit pattern, weighted by the unrotated
// weights mu (Compressor.cpp:334-339) and summed in channel order.
template <bool NU = false>
__device__ __forceinline__ uint32_t single_color(const Fit &f, const uint32_t *sctab, uint32_t pixel,
                                 uint32_t &p1, uint32_t &p2, int &pbc,
                                 const Weights *mu = nullptr) {
  uint32_t best = NU ? __float_as_uint(FLT_MAX) : 0xFFFFFFFFu;
  const int bpi = f.nbits_idx - 1;
  for (int pbi = 0; pbi < f.npbc; pbi++) {
    uint32_t err = 0, q1 = 0, q2 = 0;
    float ferr = 0.0f;
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
      if constexpr (NU) {
        const float e = (float)dist * mu->w[ci];
        ferr += e * e;
      } else {
        err += dist * dist;
      }
      q1 |= lo << (8 * ci);
      q2 |= hi << (8 * ci);
    }
    if constexpr (NU) err = __float_as_uint(ferr);
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
__device__ __forceinline__ void power_method(const float m[16], float b[4]) {
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
__device__ __forceinline__ void principal_axis(const uint32_t px[16], uint32_t smask, int ptmode, int n,
                               const float avg[4], float axis[4]) {
  // unique points in first-occurrence order; test u2.. as they appear
  int nu = 0;
  float u0[4], dir[4];
  bool collinear = true;
  uint32_t u0p = 0;
#pragma unroll 1
  for (int i = 0; i < 16; i++) {
    if (!((smask >> i) & 1)) continue;
    const uint32_t pi = point_of(px[i], ptmode);
    bool seen = false;
#pragma unroll 1
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
  // each entry sums its products in point order
  float sum[10];
#pragma unroll
  for (int e = 0; e < 10; e++) sum[e] = 0.0f;
#pragma unroll 1
  for (int p = 0; p < 16; p++) {
    if (!((smask >> p) & 1)) continue;
    const uint32_t pt = point_of(px[p], ptmode);
    float t[4];
#pragma unroll
    for (int k = 0; k < 4; k++) t[k] = (float)chan(pt, k) - avg[k];
    int e = 0;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j <= i; j++) sum[e++] += t[i] * t[j];
  }
  float cov[16];
  {
    int e = 0;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j <= i; j++) {
        // x / 3.0f for a float x as a double product, correctly rounded:
        // x / 3 is a float or, with 3 left in its denominator, at least
        // 2^-27 (relative) from every float rounding boundary
        cov[i * 4 + j] = (float)((double)sum[e++] * (1.0 / 3.0));
        cov[j * 4 + i] = cov[i * 4 + j];
      }
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

// ---------------------------------------------------------------------------
// Per-subset fit set-up: CompressCluster up to the annealing loop.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void set_fit_attr(Fit &f, const ModeAttr a, int mode, int rot,
                                             int im) {
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

__device__ __forceinline__ void set_fit(Fit &f, int mode, int rot, int im) {
  set_fit_attr(f, kModes[mode], mode, rot, im);
}

// Where a subset fit stands once CompressCluster reaches
// OptimizeEndpointsForCluster (Compressor.cpp:550-580), or its final result
// when no annealing follows (single colour, quality 0, zero error).
struct SubStart {
  uint32_t err;  // QuantizedError of the start endpoints (or the final error)
  bool sc;       // CompressSingleColor's result: the subset's error is n * err
  uint32_t p1, p2;  // endpoint state (integers): annealing start, or final endpoints
  uint64_t idx;     // indices at pixel positions of this subset
  int pbc;
  bool anneal;
};

// Index of the centroid nearest to v (squared distance, first minimum).
template <int NB>
__device__ __forceinline__ int nearest_centroid(const float c[NB][4], const float v[4]) {
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
  return minb;
}

// NU: the nonuniform metric, mu its unrotated weights (single colour), mr
// the weights of this candidate's rotation (QuantizedError).
// km: this lane's k-means accumulators in LDS (word (j * 5 + k) * 64 of
// bucket j, k = 4 the count).
template <int NB, bool NU = false>
__device__ __forceinline__ void kmeans_setup(const Fit &f, uint32_t q, const uint32_t *sctab,
                             const uint32_t px[16], uint32_t smask, int ptmode, int n,
                             const float avg[4], SubStart &out,
                             const Weights *mu, const Weights *mr, float *km) {
  float axis[4];
  principal_axis(px, smask, ptmode, n, avg, axis);
  float mindp = FLT_MAX, maxdp = -FLT_MAX;
#pragma unroll 1
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
  // Lloyd k-means from NB evenly spaced seeds (Compressor.cpp:961-1026)
  float c[NB][4];
#pragma unroll
  for (int i = 0; i < NB; i++) {
    const float s = (float)i / (float)(NB - 1);
#pragma unroll
    for (int k = 0; k < 4; k++) c[i][k] = p1[k] * (1.0f - s) + p2[k] * s;
  }
  uint32_t cnt[NB];
  bool fixed = false;
  int guard = 0;
  while (!fixed && guard++ < 4096) {
    // every point into its nearest centroid's LDS sums, in point order, so
    // each bucket adds its points in the order the reference's per-bucket
    // loop does (Compressor.cpp:997-1009)
#pragma unroll
    for (int j = 0; j < NB * 5; j++) km[j * 64] = 0.0f;
#pragma unroll 1
    for (int i = 0; i < 16; i++) {
      if (!((smask >> i) & 1)) continue;
      float v[4];
      unpack4f(point_of(px[i], ptmode), v);
      float *b = km + nearest_centroid<NB>(c, v) * 5 * 64;
#pragma unroll
      for (int k = 0; k < 4; k++) b[k * 64] += v[k];
      b[4 * 64] += 1.0f;
    }
    fixed = true;
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const float cn = km[(j * 5 + 4) * 64];
      float s[4];
      // sum / cn in float, correctly rounded, as a double product: the sum
      // is an integer below 2^13 and cn one of 1..16, so sum / cn is either
      // a float or, in lowest terms p / q with q odd, at least 2^-29
      // (relative) from every float rounding boundary, far beyond the
      // product's 2^-52
      const double inv = kInvCount[(int)cn];
#pragma unroll
      for (int k = 0; k < 4; k++)
        s[k] = cn != 0.0f ? (float)((double)km[(j * 5 + k) * 64] * inv) : 0.0f;
      if (s[0] != c[j][0] || s[1] != c[j][1] || s[2] != c[j][2] || s[3] != c[j][3])
        fixed = false;
#pragma unroll
      for (int k = 0; k < 4; k++) c[j][k] = s[k];
      cnt[j] = (uint32_t)cn;
    }
  }
  int nfilled = 0, last = 0;
#pragma unroll
  for (int j = 0; j < NB; j++)
    if (cnt[j] > 0) {
      nfilled++;
      last = j;
    }
  uint64_t ones = 0;
#pragma unroll
  for (int i = 0; i < 16; i++)
    if ((smask >> i) & 1) ones |= (uint64_t)1 << (4 * i);
  if (nfilled == 1) {
    float cl[4];
#pragma unroll
    for (int j = 0; j < NB; j++)
      if (j == last) {
#pragma unroll
        for (int k = 0; k < 4; k++) cl[k] = c[j][k];
      }
    const uint32_t e = single_color<NU>(f, sctab, to_pixel_f(cl, 0xFFFFFFFFu, -1), out.p1,
                                        out.p2, out.pbc, mu);
    out.err = NU ? e : (uint32_t)n * e;
    out.sc = true;
    out.idx = ones;
    out.anneal = false;
    return;
  }
  // least squares (Compressor.cpp:1053-1077)
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
  int pbc = 0;
  uint32_t g1 = 0xFFFFFFFFu, g2 = 0xFFFFFFFFu;
  {
    const bool has_pb = f.npbc > 1;
    float mind = FLT_MAX;
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
        g1 = q1;
        g2 = q2;
        pbc = i;
      }
    }
  }
  // OptimizeEndpointsForCluster's start (Compressor.cpp:550-580): the error
  // is measured with the combo's p-bits even for modes without p-bits (it
  // always receives a combo), the state keeps the grid endpoints.
  const int nb = 1 << f.nbits_idx;
  const uint32_t q1 = to_pixel_int(g1, f.qmask, pbc_bit(f.pbit_type, pbc, 0));
  const uint32_t q2 = to_pixel_int(g2, f.qmask, pbc_bit(f.pbit_type, pbc, 1));
  if constexpr (NU) out.err = __float_as_uint(qerr_nu<1>(px, smask, ptmode, q1, q2, nb, *mr, out.idx));
  else out.err = qerr<1>(px, smask, ptmode, q1, q2, nb, out.idx);
  out.sc = false;
  out.pbc = pbc;
  if (f.pbit_type != PB_NONE) {
    out.p1 = q1;
    out.p2 = q2;
  } else {
    out.p1 = to_pixel_int(g1, f.qmask, -1);
    out.p2 = to_pixel_int(g2, f.qmask, -1);
  }
  out.anneal = q > 0 && out.err > 0;
}

// CompressCluster, RGB variant (Compressor.cpp:921-1094) up to the annealing.
template <bool NU = false>
__device__ __forceinline__ void subset_setup(const Fit &f, uint32_t q, const uint32_t *sctab,
                             const uint32_t px[16], uint32_t smask, int ptmode, int n,
                             const float avg[4], bool all_same, SubStart &out,
                             const Weights *mu, const Weights *mr, float *km) {
  if (all_same) {
    const int first = __ffs(smask) - 1;
    uint32_t pt = 0;
#pragma unroll
    for (int i = 0; i < 16; i++)
      if (i == first) pt = point_of(px[i], ptmode);
    const uint32_t e = single_color<NU>(f, sctab, pt, out.p1, out.p2, out.pbc, mu);
    out.err = NU ? e : (uint32_t)n * e;
    out.sc = true;
    out.idx = 0;
#pragma unroll
    for (int i = 0; i < 16; i++)
      if ((smask >> i) & 1) out.idx |= (uint64_t)1 << (4 * i);
    out.anneal = false;
    return;
  }
  if (f.nbits_idx == 2)
    kmeans_setup<4, NU>(f, q, sctab, px, smask, ptmode, n, avg, out, mu, mr, km);
  else if (f.nbits_idx == 3)
    kmeans_setup<8, NU>(f, q, sctab, px, smask, ptmode, n, avg, out, mu, mr, km);
  else
    kmeans_setup<16, NU>(f, q, sctab, px, smask, ptmode, n, avg, out, mu, mr, km);
}

// The alpha channel of the rotation modes (Compressor.cpp:697-911): a 1-D
// k-means and least-squares fit (or the table lookup for a constant alpha).
// Returns the alpha error; a1q/a2q are the alpha bytes Pack will write
// (rounded, then quantised to the mode's alpha precision).  NU: the
// nonuniform metric, whose alpha term (a double sum of float per-pixel
// errors, weighted by the rotated alpha weight aw, Compressor.cpp:712) goes
// to *nu_err.
template <bool NU = false>
__device__ __forceinline__ uint32_t alpha_fit(const Fit &f, const uint32_t px[16], uint64_t &aidx, uint32_t &a1q,
                              uint32_t &a2q, float aw = 1.0f, double *nu_err = nullptr) {
  float avals[16];
  float amin = FLT_MAX, amax = -FLT_MAX;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const uint32_t src = f.rot == 0 ? 3 : (uint32_t)(f.rot - 1);
    avals[i] = (float)chan(px[i], (int)src);
    amin = avals[i] < amin ? avals[i] : amin;
    amax = amax < avals[i] ? avals[i] : amax;
  }
  const int abits = f.idx_mode == 0 ? kModes[f.mode].alpha_idx_bits : kModes[f.mode].idx_bits;
  const int nb = 1 << abits;
  float a1 = amin, a2 = amax;
  uint32_t alpha_err = 0;
  double alpha_nu = 0.0;
  aidx = 0;
  if (a1 == a2) {
    const uint32_t a1be = (uint32_t)a1, a2be = (uint32_t)a2;
    if (f.mode != 5) {
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
      if constexpr (NU) {
        float pe = aw * (float)d;
        pe *= pe;
        alpha_nu = (double)(16.0f * pe);
      }
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
        // integer sum over a count of 1..16: the double product is the
        // correctly rounded quotient (as for the colour k-means)
        if (c > 0.0f) s = (float)((double)s * kInvCount[(int)c]);
        npts[j] = c;
        if (j < nb) fixed = fixed && (s == vals[j]);
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
    // QuantizeChannel(uint8(a)) truncates here, while Pack rounds (:876-879)
    const uint32_t amask = (uint32_t)((int8_t)-0x7F >> (f.alpha_prec - 1)) & 0xFFu;
    const uint32_t a1b = quant_channel((uint32_t)a1 & 0xFFu, amask, -1);
    const uint32_t a2b = quant_channel((uint32_t)a2 & 0xFFu, amask, -1);
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const uint32_t val = (uint32_t)avals[i];
      uint32_t me = 0xFFFFFFFFu;
      float mef = FLT_MAX;
      uint32_t bb = 0;
#pragma unroll
      for (int j = 0; j < 8; j++) {
        if (j >= nb) continue;
        const int w1 = kW[abits - 1][j], w0 = 64 - w1;
        const uint32_t ip = ((a1b * w0) + (a2b * w1) + 32) >> 6 & 0xFFu;
        const uint32_t d = val > ip ? val - ip : ip - val;
        if constexpr (NU) {
          float pe = aw * (float)d;
          pe *= pe;
          if (pe < mef) {
            mef = pe;
            bb = (uint32_t)j;
          }
        } else {
          const uint32_t e = d * d;
          if (e < me) {
            me = e;
            bb = (uint32_t)j;
          }
        }
      }
      if constexpr (NU) alpha_nu += (double)mef;
      else alpha_err += me;
      aidx |= (uint64_t)bb << (4 * i);
    }
  }
  const uint32_t am = chan(f.qmask, 3);
  a1q = quant_channel(round_half_up(a1) & 0xFFu, am, -1);
  a2q = quant_channel(round_half_up(a2) & 0xFFu, am, -1);
  if constexpr (NU) *nu_err = alpha_nu;
  return alpha_err;
}

// ---------------------------------------------------------------------------
// Work items.
// ---------------------------------------------------------------------------

// Candidates of a block: one per (mode, shape entry) of CompressClusters
// (Compressor.cpp:1771-1809), with the rotation modes split per (rotation,
// index mode), stored in the order the reference tries them (modes 0, 2, 1,
// 3, 7, 4, 5, 6; shape entries; rotations; index modes), so that the first
// strict minimum over the slots is the reference's choice.  Up to 32 slots
// per block (16 in Compress's search; CompressWithStats runs modes 4, 5 and
// 6 once per shape entry), each a 16-bit descriptor:
//   mode 3 | shape 6 << 3 | rotation 2 << 9 | index mode 1 << 11 | entry 3 << 12
// (entry = the shape selection's entry, 0..4, which keys the annealing stream).
// Candidate types (work queues, sorted so that a wave runs one code path):
//   0 mode 0, 1 mode 2, 2 mode 1, 3 mode 3, 4 mode 7, 5 mode 6, 6 mode 4, 7 mode 5.
constexpr int kNumTypes = 8;
constexpr int kSlotBits = 5;
constexpr int kMaxCands = 1 << kSlotBits;
constexpr int kMaxBoxCands = 16;  // slots of a BoxSelection or learned block
// most candidates of one type a block can have (mode 4: 8 per shape entry)
constexpr int kMaxPerType = 16;

// Candidate id: relative block << kSlotBits | slot; also the candidate's
// index in the per-slot arrays (block * kMaxCands + slot).
__host__ __device__ __forceinline__ uint32_t cand_id(uint32_t rel, uint32_t slot) {
  return (rel << kSlotBits) | slot;
}
__host__ __device__ __forceinline__ uint32_t id_block(uint32_t id) { return id >> kSlotBits; }
__host__ __device__ __forceinline__ uint32_t id_slot(uint32_t id) {
  return id & (kMaxCands - 1u);
}
__constant__ int kTypeOfMode[8] = {0, 2, 1, 3, 6, 7, 5, 4};

// Block class and candidate count, packed: cls 2 | count 5 << 2 (at most 31).
// kSelFull marks a block of the learned search that runs the full search
// (the ranker was not confident): every candidate annealed, the reference's
// choice by internal error, the quality guard.
__device__ __forceinline__ uint32_t sel_pack(int cls, int count) {
  return (uint32_t)cls | ((uint32_t)count << 2);
}
constexpr uint32_t kSelFull = 1u << 7;
__device__ __forceinline__ int sel_cls(uint32_t s) { return s & 3; }
__device__ __forceinline__ int sel_count(uint32_t s) { return (s >> 2) & 31; }
__device__ __forceinline__ bool sel_full(uint32_t s) { return (s & kSelFull) != 0; }

struct Cand {
  int mode, entry, rot, im, shape, nsub;
};

__device__ __forceinline__ uint32_t cdesc_pack(int mode, int shape, int rot, int im, int entry) {
  return (uint32_t)mode | ((uint32_t)shape << 3) | ((uint32_t)rot << 9) | ((uint32_t)im << 11) |
         ((uint32_t)entry << 12);
}

__device__ __forceinline__ void cand_from_desc(uint32_t d, Cand &k) {
  k.mode = d & 7;
  k.shape = (d >> 3) & 63;
  k.rot = (d >> 9) & 3;
  k.im = (d >> 11) & 1;
  k.entry = (d >> 12) & 7;
  k.nsub = kModes[k.mode].subsets;
}

// The candidate list of a searched block after BoxSelection, in the
// reference's order (see above).  Returns the count.
__device__ int box_candidates(bool opaque_layout, int nshapes, int s2, int s3, uint32_t modes,
                              uint32_t q, uint16_t out[kMaxBoxCands]) {
  int n = 0;
  if (opaque_layout) {
    if (nshapes >= 2) {
      if (((modes >> 0) & 1) && s3 < 16) out[n++] = (uint16_t)cdesc_pack(0, s3, 0, 0, 1);
      if ((modes >> 2) & 1) out[n++] = (uint16_t)cdesc_pack(2, s3, 0, 0, 1);
    }
    if ((modes >> 1) & 1) out[n++] = (uint16_t)cdesc_pack(1, s2, 0, 0, 0);
    if ((modes >> 3) & 1) out[n++] = (uint16_t)cdesc_pack(3, s2, 0, 0, 0);
    if ((modes >> 7) & 1) out[n++] = (uint16_t)cdesc_pack(7, s2, 0, 0, 0);
    if ((modes >> 6) & 1) {
      out[n++] = (uint16_t)cdesc_pack(6, s2, 0, 0, 0);
      // both mode-6 entries fit the same 16 pixels; without annealing the
      // second repeats the first exactly and cannot win (strict <)
      if (nshapes >= 2 && q > 0) out[n++] = (uint16_t)cdesc_pack(6, s3, 0, 0, 1);
    }
  } else {
    if ((modes >> 7) & 1) out[n++] = (uint16_t)cdesc_pack(7, s2, 0, 0, 0);
    if ((modes >> 4) & 1)
      for (int c = 0; c < 8; c++) out[n++] = (uint16_t)cdesc_pack(4, 0, c >> 1, c & 1, 0);
    if ((modes >> 5) & 1)
      for (int c = 0; c < 4; c++) out[n++] = (uint16_t)cdesc_pack(5, 0, c, 0, 0);
    if ((modes >> 6) & 1) out[n++] = (uint16_t)cdesc_pack(6, s2, 0, 0, 0);
  }
  return n;
}

// CandRec::meta bit set when annealing rewrote the record.
constexpr uint32_t kMetaAnnealed = 1u << 12;

// Result of one candidate, read by the pack kernel.
struct CandRec {
  uint32_t err;       // total error; 0xFFFFFFFF = not a candidate
  uint32_t px1[3], px2[3];
  uint32_t cidx_lo, cidx_hi;
  uint32_t aidx_lo, aidx_hi;
  uint32_t meta;      // pbc per subset (2 bits each) | rot << 8 | im << 10 | annealed << 12
};

// Annealing work item (start state from the set-up kernel).
struct AState {
  uint32_t id;        // cand_id(relative block, slot)
  uint32_t info;      // mode | rot << 3 | im << 5 | nsub << 6 | shape << 8 | active << 14 | entry 3 << 17
  uint32_t p1[3], p2[3];
  uint32_t err[3];
  uint32_t pbc;       // 2 bits per subset
  uint32_t idx_lo, idx_hi;
  uint32_t extra_err; // alpha error of rotation candidates
  uint32_t pad;
};

// Learned search: candidates whose quality-0 block decodes to more than
// 1 + 1/kAnnealMarginDiv times the block's best quality-0 error are not
// annealed (annealing lowers a fit's error by a few percent at most).
constexpr unsigned long long kAnnealMarginDiv = 50;

// Learned search: the list of annealing states that pass the margin, in
// astate order (the set-up kernels append by mode, so annealing waves stay
// uniform and the list's end holds the rotation modes).  An ordered
// compaction over chunks of 4096 states: per-chunk counts, their exclusive
// scan, then each chunk writes its passing states from its offset.
constexpr uint32_t kListChunk = 4096;

// Flags of the states cb + 256 k + t (k = 0..15) of thread t: bit k set when
// the state passes the margin.  Adjacent lanes read adjacent states.
__device__ __forceinline__ uint32_t list_flags(const uint32_t *aid, uint32_t total,
                                               const uint32_t *sel,
                                               const unsigned long long *cerr,
                                               const unsigned long long *q0best, uint32_t cb,
                                               int t) {
  uint32_t flags = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const uint32_t i = cb + 256u * (uint32_t)k + (uint32_t)t;
    if (i < total) {
      const uint32_t id = aid[i];
      const uint32_t rel = id_block(id);
      const unsigned long long best = q0best[rel];
      // a full-search block anneals every candidate
      if (sel_full(sel[rel]) || cerr[id] <= best + best / kAnnealMarginDiv) flags |= 1u << k;
    }
  }
  return flags;
}

// Exclusive prefix over the 256 threads of a workgroup; returns the total.
__device__ __forceinline__ uint32_t wg_exclusive_scan(uint32_t v, uint32_t *excl,
                                                      uint32_t *s_wave) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint32_t incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(incl, o);
    if (lane >= o) incl += u;
  }
  if (lane == 63) s_wave[wave] = incl;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    before += q < wave ? s_wave[q] : 0u;
    all += s_wave[q];
  }
  *excl = before + incl - v;
  return all;
}

__global__ void __launch_bounds__(256) list_count_kernel(const uint32_t *aid, const uint32_t *counters,
                                                         const uint32_t *sel,
                                                         const unsigned long long *cerr,
                                                         const unsigned long long *q0best,
                                                         uint32_t *chunk_count) {
  __shared__ uint32_t s_wave[4];
  const uint32_t total = counters[16];
  const uint32_t cb = blockIdx.x * kListChunk;
  if (cb >= total) return;  // whole workgroup
  const uint32_t flags = list_flags(aid, total, sel, cerr, q0best, cb, threadIdx.x);
  uint32_t excl;
  const uint32_t all = wg_exclusive_scan((uint32_t)__popc(flags), &excl, s_wave);
  if (threadIdx.x == 0) chunk_count[blockIdx.x] = all;
}

// One workgroup: exclusive scan of the chunk counts (at most 64 per thread),
// the list length into counters[18].
__global__ void __launch_bounds__(256) list_scan_kernel(uint32_t *counters, uint32_t *chunk_count) {
  __shared__ uint32_t s_wave[4];
  const uint32_t nchunks = (counters[16] + kListChunk - 1) / kListChunk;
  const uint32_t per = (nchunks + 255) / 256;
  const uint32_t c0 = threadIdx.x * per;
  uint32_t sum = 0;
  for (uint32_t c = c0; c < c0 + per && c < nchunks; c++) sum += chunk_count[c];
  uint32_t excl;
  const uint32_t all = wg_exclusive_scan(sum, &excl, s_wave);
  for (uint32_t c = c0; c < c0 + per && c < nchunks; c++) {
    const uint32_t v = chunk_count[c];
    chunk_count[c] = excl;
    excl += v;
  }
  if (threadIdx.x == 0) counters[18] = all;
}

__global__ void __launch_bounds__(256) list_write_kernel(const uint32_t *aid, const uint32_t *counters,
                                                         const uint32_t *sel,
                                                         const unsigned long long *cerr,
                                                         const unsigned long long *q0best,
                                                         const uint32_t *chunk_offset,
                                                         uint32_t *alist) {
  const uint32_t total = counters[16];
  const uint32_t cb = blockIdx.x * kListChunk;
  if (cb >= total) return;  // whole workgroup
  const uint32_t flags = list_flags(aid, total, sel, cerr, q0best, cb, threadIdx.x);
  // round k places the passing states cb + 256 k .. cb + 256 k + 255 in
  // order: rank within the wave by ballot, then the waves before it
  __shared__ uint32_t s_round[2][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t pos = chunk_offset[blockIdx.x];
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const bool f = (flags >> k) & 1u;
    const uint64_t m = __ballot(f);
    if (lane == 0) s_round[k & 1][wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      before += q < wave ? s_round[k & 1][q] : 0u;
      all += s_round[k & 1][q];
    }
    if (f)
      alist[pos + before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] =
          cb + 256u * (uint32_t)k + (uint32_t)threadIdx.x;
    pos += all;
  }
}

struct Work {
  uint32_t *sel;          // per block
  uint16_t *cdesc;        // per block x 16 candidate descriptors
  CandRec *rec;           // per block x 16
  uint32_t *queue;        // per-type candidate queues, kMaxPerType entries per block each
  uint32_t *counters;     // [0..7] queue sizes, [8..15] set-up fetch, [16] anneal size,
                          // [17] anneal fetch, [18] learned annealing list size
  AState *astate;
  uint32_t qcap;          // entries per type queue
  uint32_t astate_cap;
  // learned search: the annealing states to run (indices into astate, in
  // astate order), counters[18] of them; null in the reference search
  const uint32_t *alist;
  // learned search: astate[i].id for every state (read when the list is
  // built, without touching the 64-byte states); null in the reference search
  uint32_t *aid;
  // nonuniform metric: each candidate's total error, the double the
  // reference compares (per block x kMaxCands); null under the uniform metric
  double *nu_total;
};

__device__ __forceinline__ void load_block(const BlockRange &r, uint32_t rel, uint32_t px[16]) {
  const uint32_t b = r.first_block + rel;
  const uint32_t bx = b % r.blocks_wide, by = b / r.blocks_wide;
  const uint32_t *p = r.img + (size_t)(by * 4) * r.width + bx * 4;
#pragma unroll
  for (int y = 0; y < 4; y++) {
    const uint4 row = *reinterpret_cast<const uint4 *>(p + (size_t)y * r.width);
    px[y * 4 + 0] = row.x;
    px[y * 4 + 1] = row.y;
    px[y * 4 + 2] = row.z;
    px[y * 4 + 3] = row.w;
  }
}

// Key of one fit's annealing stream; shape entries 0..4 take bits 3-4 and 10.
__device__ __forceinline__ uint32_t cand_tag(int mode, int entry, int subset, int rot, int im) {
  return (uint32_t)mode | ((uint32_t)(entry & 3) << 3) | ((uint32_t)subset << 5) |
         ((uint32_t)rot << 7) | ((uint32_t)im << 9) | ((uint32_t)(entry >> 2) << 10);
}

// Appends the workgroup's valid candidates to their type queues: slots are
// taken with LDS atomics, then one global atomic per type per workgroup.
// Every thread of the workgroup must call it.
__device__ __forceinline__ void wg_enqueue(Work &w, bool valid, int type, uint32_t id,
                                           uint32_t *s_cnt, uint32_t *s_base) {
  const int t = threadIdx.x;
  if (t < kNumTypes) s_cnt[t] = 0;
  __syncthreads();
  uint32_t local = 0;
  if (valid) local = atomicAdd(&s_cnt[type], 1u);
  __syncthreads();
  if (t < kNumTypes && s_cnt[t]) s_base[t] = atomicAdd(&w.counters[t], s_cnt[t]);
  __syncthreads();
  if (valid) w.queue[(size_t)type * w.qcap + s_base[type] + local] = id;
}

// Packed 16-bit lanes: channels 0 and 2 of a pixel are (pix & 0x00FF00FF),
// channels 1 and 3 are ((pix >> 8) & 0x00FF00FF); every quantity of the
// index search fits 16 bits (interpolation sums <= 255 * 64 + 32, channel
// differences within +-255), so v_pk_* and v_dot2 give the exact integers.
typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ u16x2 as_u16x2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
__device__ __forceinline__ s16x2 as_s16x2(uint32_t v) { return __builtin_bit_cast(s16x2, v); }
__device__ __forceinline__ s16x2 as_s16x2(u16x2 v) { return __builtin_bit_cast(s16x2, v); }

// Weight of the second endpoint at index j for nb = 4, 8, 16 entries:
// (64 j + (nb - 1) / 2) / (nb - 1) by a multiply-shift, equal to
// kInterpolationValues (Compressor.cpp:159-168); bc7_upload_tables checks
// every entry on the host before the kernels run.
struct WeightFn {
  uint32_t h, m;
};
__host__ __device__ __forceinline__ WeightFn weight_fn(int nb) {
  if (nb == 4) return {1u, 21846u};
  if (nb == 8) return {3u, 9363u};
  return {7u, 4370u};
}
__host__ __device__ __forceinline__ uint32_t weight_of(WeightFn fn, uint32_t j) {
  // j < 16: with the mask the compiler sees both factors below 2^24 and
  // uses the full-rate 24-bit multiply
  return ((64u * (j & 15u) + fn.h) * fn.m) >> 16;
}

// Per-subset constants of one annealing step.  The interpolated value is
// (e1 * (64 - w) + e2 * w + 32) >> 6 = (64 e1 + 32 + (e2 - e1) * w) >> 6, so
// 64 e1 + 32 and the difference suffice.
struct SubConst {
  uint32_t a02, a13;  // 64 * endpoint 1 + 32, channel pairs (see interp_err)
  uint32_t d02, d13;  // endpoint 2 - endpoint 1 (signed pairs)
  int ad;             // endpoint 1 . (endpoint 2 - endpoint 1)
  int l;              // squared endpoint distance, 1 when the endpoints are equal
  float rl;           // 1 / l from below (see pixel_bounds)
};

__device__ __forceinline__ SubConst sub_const(uint32_t e1, uint32_t e2) {
  SubConst c;
  const uint32_t a02 = e1 & 0x00FF00FFu, a13 = (e1 >> 8) & 0x00FF00FFu;
  const uint32_t b02 = e2 & 0x00FF00FFu, b13 = (e2 >> 8) & 0x00FF00FFu;
  c.d02 = __builtin_bit_cast(uint32_t, as_s16x2(b02) - as_s16x2(a02));
  c.d13 = __builtin_bit_cast(uint32_t, as_s16x2(b13) - as_s16x2(a13));
  // each 16-bit lane: 64 a + 32 <= 16352
  c.a02 = (a02 << 6) + 0x00200020u;
  c.a13 = (a13 << 6) + 0x00200020u;
  // endpoint 1's projection on the axis, subtracted from each pixel's
  c.ad = __builtin_amdgcn_sdot2(as_s16x2(a02), as_s16x2(c.d02),
                                __builtin_amdgcn_sdot2(as_s16x2(a13), as_s16x2(c.d13), 0, false),
                                false);
  const int l = __builtin_amdgcn_sdot2(as_s16x2(c.d02), as_s16x2(c.d02),
                                       __builtin_amdgcn_sdot2(as_s16x2(c.d13),
                                                              as_s16x2(c.d13), 0, false),
                                       false);
  // equal endpoints give d = 0 for every pixel, so l = 1 yields index 0
  c.l = l > 0 ? l : 1;
  // v_rcp_f32 is within 1 ulp (2^-23); scaled by 1 - 2^-20 the product
  // num * rl stays strictly below num / l for num > 0
  c.rl = __builtin_amdgcn_rcpf((float)c.l) * (1.0f - 0x1p-20f);
  return c;
}

// Squared error of the pixel pairs against index weight w1: each lane
// computes (64 e1 + 32 + (e2 - e1) w1) >> 6, where (e2 - e1) w1 lies within
// +-255 * 64 and the sum is 64 times an interpolant in 0..255 plus 32 (at
// most 16352), so signed 16-bit lanes hold every intermediate exactly.
__device__ __forceinline__ uint32_t interp_err(uint32_t p02, uint32_t p13, const SubConst &c,
                                               uint32_t w1) {
  const s16x2 W = {(short)w1, (short)w1};
  const s16x2 S = {6, 6};
  const s16x2 i02 = (as_s16x2(c.d02) * W + as_s16x2(c.a02)) >> S;
  const s16x2 i13 = (as_s16x2(c.d13) * W + as_s16x2(c.a13)) >> S;
  const s16x2 e02 = as_s16x2(p02) - i02;
  const s16x2 e13 = as_s16x2(p13) - i13;
  return (uint32_t)__builtin_amdgcn_sdot2(e02, e02, __builtin_amdgcn_sdot2(e13, e13, 0, false),
                                          false);
}

// QuantizedError's per-pixel index search (RGBAEndpoints.cpp:240-300) for
// pixel pix (projected as pt), without branches: both neighbouring indices
// are scored and the second kept only when it exists and is strictly better;
// equal endpoints give index 0.  Returns the error, the index in *best.
//
// The reference takes j1 = floor(t), j2 = ceil(t) of t = fl(fl(d / l) * kf)
// (float, d the integer projection, l the integer squared distance, kf =
// nb - 1).  Those equal the floor and ceiling of the rational d kf / l: its
// distance to an integer, when nonzero, is at least 1 / l >= 1 / (4 * 255^2)
// = 3.8e-6, more than the float error of at most 2^-23 |t| <= 1.9e-6 for
// |t| <= 16 (larger |t| is clamped away), and fl(fl(k / kf) kf) = k for every
// integer 0 <= k <= kf in {3, 7, 15} (checked exhaustively).  So the quotient
// is taken in integers: an estimate from a reciprocal, corrected by one step
// with the exact remainder.  The estimate comes from below (rl < 1 / l with
// a margin larger than the float errors, and num / l <= 15 * 510 / sqrt(l)
// keeps its shortfall under 1), so for num >= 0 it is the quotient or one
// less, and one upward step fixes it.  For num < 0 the index is 0 whatever
// the estimate: j1 clamps to 0 and j2 <= j1.
__device__ __forceinline__ void pixel_bounds(uint32_t pt, const SubConst &c, int nb, int &j1,
                                             int &j2) {
  const uint32_t t02 = pt & 0x00FF00FFu, t13 = (pt >> 8) & 0x00FF00FFu;
  const int d = __builtin_amdgcn_sdot2(
      as_s16x2(t02), as_s16x2(c.d02),
      __builtin_amdgcn_sdot2(as_s16x2(t13), as_s16x2(c.d13), -c.ad, false),
      false);
  const int num = __mul24(d, nb - 1);  // |d| <= 4 * 255^2, so 24-bit operands suffice
  int q = (int)floorf((float)num * c.rl);
  int rem = num - __mul24(q, c.l);
  const bool hi = rem >= c.l;
  q += hi ? 1 : 0;
  rem -= hi ? c.l : 0;
  j1 = q < 0 ? 0 : (q > nb - 1 ? nb - 1 : q);
  j2 = q + (rem != 0 ? 1 : 0);
  j2 = j2 > nb - 1 ? nb - 1 : j2;
}

__device__ __forceinline__ uint32_t pixel_err(uint32_t pix, uint32_t pt, const SubConst &c,
                                              int nb, float kf, WeightFn wf, int *best) {
  (void)kf;
  const uint32_t p02 = pix & 0x00FF00FFu, p13 = (pix >> 8) & 0x00FF00FFu;
  int j1, j2;
  pixel_bounds(pt, c, nb, j1, j2);
  const uint32_t e1 = interp_err(p02, p13, c, weight_of(wf, (uint32_t)j1));
  const uint32_t e2 = interp_err(p02, p13, c, weight_of(wf, (uint32_t)j2));
  const bool second = j2 > j1 && e2 < e1;
  *best = second ? j2 : j1;
  return second ? e2 : e1;
}

// The nonuniform metric's pixel pass (the metric is described above qerr_nu).
// interp_err under the weights (pair lanes: .x channels 0 / 1, .y 2 / 3).
__device__ __forceinline__ float interp_err_nu(uint32_t p02, uint32_t p13, const SubConst &c,
                                               uint32_t w1, const Weights &m) {
  const s16x2 W = {(short)w1, (short)w1};
  const s16x2 S = {6, 6};
  const s16x2 i02 = (as_s16x2(c.d02) * W + as_s16x2(c.a02)) >> S;
  const s16x2 i13 = (as_s16x2(c.d13) * W + as_s16x2(c.a13)) >> S;
  const s16x2 e02 = as_s16x2(p02) - i02;
  const s16x2 e13 = as_s16x2(p13) - i13;
  return werr(e02.x, e13.x, e02.y, e13.y, m);
}

// pixel_err under the weights: the same two neighbouring indices, the second
// kept only when strictly better.
__device__ __forceinline__ float pixel_err_nu(uint32_t pix, uint32_t pt, const SubConst &c,
                                              int nb, WeightFn wf, const Weights &m,
                                              int *best) {
  const uint32_t p02 = pix & 0x00FF00FFu, p13 = (pix >> 8) & 0x00FF00FFu;
  int j1, j2;
  pixel_bounds(pt, c, nb, j1, j2);
  const float e1 = interp_err_nu(p02, p13, c, weight_of(wf, (uint32_t)j1), m);
  const float e2 = interp_err_nu(p02, p13, c, weight_of(wf, (uint32_t)j2), m);
  const bool second = j2 > j1 && e2 < e1;
  *best = second ? j2 : j1;
  return second ? e2 : e1;
}

__device__ __forceinline__ SubConst pick_sub(bool c, const SubConst &x, const SubConst &y) {
  SubConst r;
  r.a02 = c ? x.a02 : y.a02;
  r.a13 = c ? x.a13 : y.a13;
  r.ad = c ? x.ad : y.ad;
  r.d02 = c ? x.d02 : y.d02;
  r.d13 = c ? x.d13 : y.d13;
  r.l = c ? x.l : y.l;
  r.rl = c ? x.rl : y.rl;
  return r;
}

// Subset of pixel i under shape masks m1, m2 (0 when the bit of m1 is clear).
__device__ __forceinline__ uint32_t subset_of(uint32_t m1, uint32_t m2, int i) {
  return ((m1 >> i) & 1u) ? 1u + ((m2 >> i) & 1u) : 0u;
}

// Bounding-box corners of every subset of a shape (GetBoundingBox after
// SetPartition, RGBAEndpoints.h:205-226), as packed RGBA8.
template <int NS>
__device__ __forceinline__ void subset_corners(const uint32_t px[16], uint32_t m1, uint32_t m2,
                                               uint32_t lo[NS], uint32_t hi[NS]) {
  uint32_t mn02[NS], mn13[NS], mx02[NS], mx13[NS];
#pragma unroll
  for (int t = 0; t < NS; t++) {
    mn02[t] = mn13[t] = 0x00FF00FFu;
    mx02[t] = mx13[t] = 0u;
  }
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const uint32_t s = subset_of(m1, m2, i);
    const u16x2 p02 = as_u16x2(px[i] & 0x00FF00FFu), p13 = as_u16x2((px[i] >> 8) & 0x00FF00FFu);
#pragma unroll
    for (int t = 0; t < NS; t++) {
      const bool in = s == (uint32_t)t;
      const u16x2 top = {0xFF, 0xFF}, bottom = {0, 0};
      mn02[t] = __builtin_bit_cast(uint32_t,
                                   __builtin_elementwise_min(as_u16x2(mn02[t]), in ? p02 : top));
      mn13[t] = __builtin_bit_cast(uint32_t,
                                   __builtin_elementwise_min(as_u16x2(mn13[t]), in ? p13 : top));
      mx02[t] = __builtin_bit_cast(uint32_t,
                                   __builtin_elementwise_max(as_u16x2(mx02[t]), in ? p02 : bottom));
      mx13[t] = __builtin_bit_cast(uint32_t,
                                   __builtin_elementwise_max(as_u16x2(mx13[t]), in ? p13 : bottom));
    }
  }
#pragma unroll
  for (int t = 0; t < NS; t++) {
    lo[t] = mn02[t] | (mn13[t] << 8);
    hi[t] = mx02[t] | (mx13[t] << 8);
  }
}

// BoxSelection's estimate of one partition shape (EstimateTwoClusterError /
// EstimateThreeClusterError, Compressor.cpp:1626-1655): the sum over subsets
// of 0.0001 + QuantizedError of the subset's points between its bounding-box
// corners (no quantisation, NB index entries), 0 for a one-colour subset.  Each
// pixel is scored once, against its own subset's corners; per-subset errors
// are exact integers, added in double in subset order as the reference does.
template <int NS, int NB>
__device__ __forceinline__ double shape_estimate(const uint32_t px[16], int shape) {
  const uint32_t m1 = NS == 2 ? (uint32_t)kShape2[shape] : (uint32_t)kShape3[shape][0];
  const uint32_t m2 = NS == 3 ? (uint32_t)kShape3[shape][1] : 0u;
  uint32_t lo[NS], hi[NS];
  subset_corners<NS>(px, m1, m2, lo, hi);
  SubConst c[NS];
  bool flat[NS];
#pragma unroll
  for (int t = 0; t < NS; t++) {
    c[t] = sub_const(lo[t], hi[t]);
    flat[t] = lo[t] == hi[t];
  }
  const WeightFn wf = weight_fn(NB);
  uint32_t e[NS];
#pragma unroll
  for (int t = 0; t < NS; t++) e[t] = 0;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const uint32_t s = ((m1 >> i) & 1u) ? 1u + ((m2 >> i) & 1u) : 0u;
    const SubConst k = NS == 2 ? pick_sub(s != 0, c[1], c[0])
                               : pick_sub(s == 0, c[0], pick_sub(s == 1, c[1], c[NS - 1]));
    int best;
    const uint32_t me = pixel_err(px[i], px[i], k, NB, (float)(NB - 1), wf, &best);
#pragma unroll
    for (int t = 0; t < NS; t++) e[t] += s == (uint32_t)t ? me : 0u;
  }
  double err = 0.0;
#pragma unroll
  for (int t = 0; t < NS; t++) {
    if (!flat[t]) err += 0.0001 + (double)e[t];
    else err += 0.0;
  }
  return err;
}

// shape_estimate under the nonuniform metric's unrotated weights
// (Compressor.cpp:1634, 1650): each subset's error is a float sum in pixel
// order, added to 0.0001 in double.
template <int NS, int NB>
__device__ __forceinline__ double shape_estimate_nu(const uint32_t px[16], int shape,
                                                    const Weights &m) {
  const uint32_t m1 = NS == 2 ? (uint32_t)kShape2[shape] : (uint32_t)kShape3[shape][0];
  const uint32_t m2 = NS == 3 ? (uint32_t)kShape3[shape][1] : 0u;
  uint32_t lo[NS], hi[NS];
  subset_corners<NS>(px, m1, m2, lo, hi);
  SubConst c[NS];
  bool flat[NS];
#pragma unroll
  for (int t = 0; t < NS; t++) {
    c[t] = sub_const(lo[t], hi[t]);
    flat[t] = lo[t] == hi[t];
  }
  const WeightFn wf = weight_fn(NB);
  float e[NS];
#pragma unroll
  for (int t = 0; t < NS; t++) e[t] = 0.0f;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const uint32_t s = ((m1 >> i) & 1u) ? 1u + ((m2 >> i) & 1u) : 0u;
    const SubConst k = NS == 2 ? pick_sub(s != 0, c[1], c[0])
                               : pick_sub(s == 0, c[0], pick_sub(s == 1, c[1], c[NS - 1]));
    int best;
    const float me = pixel_err_nu(px[i], px[i], k, NB, wf, m, &best);
#pragma unroll
    for (int t = 0; t < NS; t++) e[t] += s == (uint32_t)t ? me : 0.0f;
  }
  double err = 0.0;
#pragma unroll
  for (int t = 0; t < NS; t++) {
    if (!flat[t]) err += 0.0001 + (double)e[t];
  }
  return err;
}

// ---------------------------------------------------------------------------
// Kernel 1: classification and BoxSelection (Compressor.cpp:1670-1750,
// 1819-1860), lane i scoring shape i; the candidates go to per-type queues.
// ---------------------------------------------------------------------------
// Classification (solid / transparent / searched) and BoxSelection of one
// block by a whole wave (lane i scores two- and three-subset shape i); the
// block's candidates in CompressClusters' order go to cd.  Returns the
// packed class and count.
template <bool NU>
__device__ uint32_t box_select_block(const uint32_t px[16], int lane, const Bc7Params &prm,
                                     uint16_t cd[kMaxBoxCands]) {
  bool solid = true, transparent = true, opaque = true;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    solid = solid && px[i] == px[0];
    transparent = transparent && (px[i] >> 24) == 0;
    opaque = opaque && (px[i] >> 24) >= 250;
  }
  uint32_t sel;
  int count = 0;
  if (solid) {
    sel = sel_pack(0, 0);
  } else if (transparent) {
    sel = sel_pack(1, 0);
  } else {
    double err2, err3;
    if constexpr (NU) {
      const Weights m = params_weights(prm);
      err2 = shape_estimate_nu<2, 8>(px, lane, m);
      err3 = opaque ? shape_estimate_nu<3, 4>(px, lane, m) : 0.0;
    } else {
      err2 = shape_estimate<2, 8>(px, lane);
      err3 = opaque ? shape_estimate<3, 4>(px, lane) : 0.0;
    }
    double v2 = err2, v3 = err3;
    int i2 = lane, i3 = lane;
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
    uint32_t modes;
    bool opaque_layout = true;
    int nshapes = 1, s3 = 0;
    if (v2 < 1e-9) {
      modes = (1u << 1) | (1u << 3) | (1u << 7);
    } else if (!opaque) {
      modes = (1u << 4) | (1u << 5) | (1u << 6) | (1u << 7);
      opaque_layout = false;
    } else {
      modes = 0xFFu & ~((1u << 4) | (1u << 5));
      nshapes = 2;
      s3 = i3;
      if (v3 < 1e-9) modes = (1u << 0) | (1u << 2);
    }
    count = box_candidates(opaque_layout, nshapes, i2, s3, modes & prm.block_modes,
                           prm.quality, cd);
    sel = sel_pack(2, count);
  }
  return sel;
}

// ---------------------------------------------------------------------------
// CompressWithStats' block path (Compressor.cpp:1954-2183): its own shape
// scan and the statistics it logs.
// ---------------------------------------------------------------------------
// QuantizedError of every subset of a shape between its bounding-box corners
// quantised with endpoint masks mask_a and mask_b (ToPixel without p-bit,
// RGBAEndpoints.cpp:167-177), NB index entries: ea/eb per subset (exact
// integers), flat when the subset is one colour.  Each pixel is scored once
// per mask, against its own subset's endpoints; TWO = false scores mask_a only.
template <int NS, int NB, bool TWO = true>
__device__ __forceinline__ void masked_subset_errors(const uint32_t px[16], uint32_t m1,
                                                     uint32_t m2, uint32_t mask_a,
                                                     uint32_t mask_b, bool flat[NS],
                                                     uint32_t ea[NS], uint32_t eb[NS]) {
  uint32_t lo[NS], hi[NS];
  subset_corners<NS>(px, m1, m2, lo, hi);
  SubConst ca[NS], cb[NS];
#pragma unroll
  for (int t = 0; t < NS; t++) {
    flat[t] = lo[t] == hi[t];
    ca[t] = sub_const(to_pixel_int(lo[t], mask_a, -1), to_pixel_int(hi[t], mask_a, -1));
    cb[t] = sub_const(to_pixel_int(lo[t], mask_b, -1), to_pixel_int(hi[t], mask_b, -1));
    ea[t] = eb[t] = 0;
  }
  const WeightFn wf = weight_fn(NB);
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const uint32_t sub = subset_of(m1, m2, i);
    SubConst ka = ca[0], kb = cb[0];
#pragma unroll
    for (int t = 1; t < NS; t++) {
      ka = pick_sub(sub == (uint32_t)t, ca[t], ka);
      kb = pick_sub(sub == (uint32_t)t, cb[t], kb);
    }
    int best;
    const uint32_t xa = pixel_err(px[i], px[i], ka, NB, (float)(NB - 1), wf, &best);
    const uint32_t xb = TWO ? pixel_err(px[i], px[i], kb, NB, (float)(NB - 1), wf, &best) : 0u;
#pragma unroll
    for (int t = 0; t < NS; t++) {
      ea[t] += sub == (uint32_t)t ? xa : 0u;
      eb[t] += sub == (uint32_t)t ? xb : 0u;
    }
  }
}

// std::min(a, b): b only when strictly smaller.
__device__ __forceinline__ double min_ref(double a, double b) { return b < a ? b : a; }

// One shape of the statistics scan (EstimateTwoClusterErrorStats and
// EstimateThreeClusterErrorStats, Compressor.cpp:1862-1935, summed over the
// subsets as at :2074-2092 and :2124-2141): the selection error and the two
// per-mode estimates (modes 1 and 3 for two subsets, 0 and 2 for three), in
// the reference's double operations and order.
template <int NS>
__device__ __forceinline__ double stats_shape(const uint32_t px[16], int shape, double &est_a,
                                              double &est_b) {
  constexpr int NB = NS == 2 ? 8 : 4;
  const uint32_t m1 = NS == 2 ? (uint32_t)kShape2[shape] : (uint32_t)kShape3[shape][0];
  const uint32_t m2 = NS == 3 ? (uint32_t)kShape3[shape][1] : 0u;
  const uint32_t mask_a = NS == 2 ? 0xFFFCFCFCu : 0xFFF0F0F0u;
  const uint32_t mask_b = NS == 2 ? 0xFFFEFEFEu : 0xFFF8F8F8u;
  bool flat[NS];
  uint32_t ea[NS], eb[NS];
  masked_subset_errors<NS, NB>(px, m1, m2, mask_a, mask_b, flat, ea, eb);
  double err = 0.0;
#pragma unroll
  for (int t = 0; t < NS; t++) {
    double x = 0.0, sa = 0.0, sb = 0.0;
    if (!flat[t]) {
      // two subsets: the estimates are the errors; three: 0.0001 + the errors
      sa = NS == 2 ? (double)ea[t] : 0.0001 + (double)ea[t];
      sb = NS == 2 ? (double)eb[t] : 0.0001 + (double)eb[t];
      x = 0.0001;
      x += min_ref(sa, sb);
    }
    err += x;
    est_a = t == 0 ? sa : est_a + sa;
    est_b = t == 0 ? sb : est_b + sb;
  }
  return err;
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmin(v, __shfl_xor(v, off));
  return v;
}

// The scan over 64 shapes, lane i holding shape i's error: the scan stops
// after the first shape below 1e-9 (brk, 64 if none); a shape is logged when
// strictly below every earlier scanned one (bestError starts at DBL_MAX).
// The logged errors go to log[], in shape order, the rest of its 64 entries
// are zeroed; returns the mask of logged shapes (the last of them is the
// chosen shape).
__device__ __forceinline__ uint64_t stats_scan(double err, int lane, int &brk, double *log) {
  const uint64_t under = __ballot(err < 1e-9);
  brk = under ? __ffsll((unsigned long long)under) - 1 : 64;
  const bool scanned = lane <= brk;
  double pm = scanned ? err : DBL_MAX;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double u = __shfl_up(pm, o);
    if (lane >= o) pm = fmin(pm, u);
  }
  double before = __shfl_up(pm, 1);
  if (lane == 0) before = DBL_MAX;
  const bool logged = scanned && err < before;
  const uint64_t mask = __ballot(logged);
  if (logged) log[__popcll(mask & ((1ull << lane) - 1ull))] = err;
  if (lane >= __popcll(mask)) log[lane] = 0.0;
  return mask;
}

// The candidates of a statistics block in CompressClusters' order
// (Compressor.cpp:1771-1809) for the shape entries (2 subsets, s2) and, on
// opaque blocks, (3 subsets, s3): a single-subset mode runs once per entry;
// without annealing its second entry repeats the first exactly and cannot
// win (strict <), so it is left out.
__device__ int stats_candidates(int nshapes, int s2, int s3, uint32_t modes, uint32_t q,
                                uint16_t out[kMaxCands]) {
  constexpr int kOrder[8] = {0, 2, 1, 3, 7, 4, 5, 6};
  int n = 0;
#pragma unroll
  for (int mi = 0; mi < 8; mi++) {
    const int mode = kOrder[mi];
    if (!((modes >> mode) & 1)) continue;
    for (int e = 0; e < nshapes; e++) {
      const int parts = e == 0 ? 2 : 3, idx = e == 0 ? s2 : s3;
      const int nsub = kModes[mode].subsets;
      if (nsub != 1 && nsub != parts) continue;
      if (mode == 0 && idx >= 16) continue;
      if (nsub == 1 && e > 0 && q == 0) continue;
      if (mode == 4) {
        for (int c = 0; c < 8; c++) out[n++] = (uint16_t)cdesc_pack(4, 0, c >> 1, c & 1, e);
      } else if (mode == 5) {
        for (int c = 0; c < 4; c++) out[n++] = (uint16_t)cdesc_pack(5, 0, c, 0, e);
      } else {
        out[n++] = (uint16_t)cdesc_pack(mode, idx, 0, 0, e);
      }
    }
  }
  return n;
}

// Classification, the statistics scan and the candidates of one block by a
// whole wave (lane i scores shape i); writes the block's statistics except
// the mode (known once the block is packed).  Returns the packed class and
// count.
__device__ uint32_t stats_select_block(const uint32_t px[16], int lane, const Bc7Params &prm,
                                       uint16_t cd[kMaxCands], fh_bc7_block_stats *bs) {
  bool solid = true, transparent = true, opaque = true;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    solid = solid && px[i] == px[0];
    transparent = transparent && (px[i] >> 24) == 0;
    opaque = opaque && (px[i] >> 24) == 255;
  }
  // every field of the record is written once (the caller's buffer may hold
  // anything): estimates -1 unless set, unused shape entries 0
  if (lane < 8) bs->mode_error[lane] = -1.0;
  if (!opaque || solid || transparent) bs->three_shape[lane] = 0.0;
  if (solid || transparent) {
    bs->two_shape[lane] = 0.0;
    if (lane < 8) bs->mode_estimate[lane] = -1.0;
    if (lane == 0) {
      bs->path = solid ? 0 : 1;
      bs->num_two_shape = bs->num_three_shape = 0;
    }
    return sel_pack(solid ? 0 : 1, 0);
  }
  // mode 6: the whole block as one line, four index values, 7-bit
  // endpoints (Compressor.cpp:2046-2061); never one colour here
  double est6;
  {
    bool flat[1];
    uint32_t ea[1], eb[1];
    masked_subset_errors<1, 4, false>(px, 0u, 0u, 0xFEFEFEFEu, 0xFEFEFEFEu, flat, ea, eb);
    est6 = flat[0] ? 0.0 : 0.0001 + (double)ea[0];
  }
  double e1, e3;
  const double err2 = stats_shape<2>(px, lane, e1, e3);
  int brk2;
  const uint64_t log2 = stats_scan(err2, lane, brk2, bs->two_shape);
  const int s2 = 63 - __clzll((long long)log2);
  uint32_t modes = 0xFFu;
  int path = 3;
  if (brk2 < 64) {
    path = 2;
    modes = (1u << 1) | (1u << 3) | (1u << 7);
  }
  // UpdateErrorEstimate keeps the minimum over the scanned shapes
  const double est1 = wave_min(lane <= brk2 ? e1 : DBL_MAX);
  const double est3 = wave_min(lane <= brk2 ? e3 : DBL_MAX);
  int s3 = 0, nshapes = 1;
  uint64_t log3 = 0;
  double est0 = -1.0, est2 = -1.0;
  if (opaque) {
    // runs even when the two-subset scan stopped early
    double e0, e2;
    const double err3 = stats_shape<3>(px, lane, e0, e2);
    int brk3;
    log3 = stats_scan(err3, lane, brk3, bs->three_shape);
    s3 = 63 - __clzll((long long)log3);
    nshapes = 2;
    if (brk3 < 64) {
      path = 2;
      modes = (1u << 0) | (1u << 2);
    }
    est0 = wave_min(lane <= brk3 ? e0 : DBL_MAX);
    est2 = wave_min(lane <= brk3 ? e2 : DBL_MAX);
  }
  if (lane < 8) {
    // modes 4, 5 and 7 get no estimate on this path
    const double est = lane == 0 ? est0 : lane == 1 ? est1 : lane == 2 ? est2
                     : lane == 3 ? est3 : lane == 6 ? est6 : -1.0;
    bs->mode_estimate[lane] = est;
  }
  if (lane == 0) {
    bs->path = path;
    bs->num_two_shape = (uint32_t)__popcll(log2);
    bs->num_three_shape = (uint32_t)__popcll(log3);
  }
  const int count = stats_candidates(nshapes, s2, s3, modes & prm.block_modes, prm.quality, cd);
  return sel_pack(2, count);
}

// The candidates of a caller's ShapeSelection (CompressionSettings::
// m_ShapeSelectionFn, Compressor.cpp:1840-1858) in CompressClusters' order
// (:1752-1809): the first min(5, num_shapes) entries, or one entry and no
// multi-subset mode when there are none.  As in the statistics search, a
// single-subset mode's entries after the first are left out without
// annealing.  At most 31 slots (the host checks every selection first).
__device__ int given_candidates(const fh_shape_selection &g, uint32_t block_modes, uint32_t q,
                                uint16_t out[kMaxCands]) {
  constexpr int kOrder[8] = {0, 2, 1, 3, 7, 4, 5, 6};
  uint32_t modes = g.selected_modes & block_modes;
  int nsh = g.num_shapes < 5 ? (int)g.num_shapes : 5;
  if (nsh == 0) {
    nsh = 1;
    modes &= ~((1u << 0) | (1u << 1) | (1u << 2) | (1u << 3) | (1u << 7));
  }
  int n = 0;
  auto emit = [&](uint32_t d) {
    if (n < kMaxCands - 1) out[n++] = (uint16_t)d;
  };
#pragma unroll
  for (int mi = 0; mi < 8; mi++) {
    const int mode = kOrder[mi];
    if (!((modes >> mode) & 1)) continue;
    for (int e = 0; e < nsh; e++) {
      const uint32_t parts = g.shapes[e].num_partitions, idx = g.shapes[e].index & 63u;
      const int nsub = kModes[mode].subsets;
      if (nsub != 1 && (uint32_t)nsub != parts) continue;
      if (mode == 0 && g.shapes[e].index >= 16) continue;
      if (nsub == 1 && e > 0 && q == 0) continue;
      if (mode == 4) {
        for (int c = 0; c < 8; c++) emit(cdesc_pack(4, 0, c >> 1, c & 1, e));
      } else if (mode == 5) {
        for (int c = 0; c < 4; c++) emit(cdesc_pack(5, 0, c, 0, e));
      } else {
        emit(cdesc_pack(mode, (int)idx, 0, 0, e));
      }
    }
  }
  return n;
}

// Classification and a caller's selection of one block (solid and
// transparent blocks never reach the selection function).
__device__ uint32_t given_select_block(const uint32_t px[16], const Bc7Params &prm,
                                       const fh_shape_selection &g, uint16_t cd[kMaxCands]) {
  bool solid = true, transparent = true;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    solid = solid && px[i] == px[0];
    transparent = transparent && (px[i] >> 24) == 0;
  }
  if (solid || transparent) return sel_pack(solid ? 0 : 1, 0);
  return sel_pack(2, given_candidates(g, prm.block_modes, prm.quality, cd));
}

// Kernel 1: one workgroup of four waves selects 16 blocks, each wave four
// in turn; the workgroup's candidates go to the type queues with one global
// atomic per type.  list, if given, holds the blocks to select (list_count
// of them, a device word); otherwise all blocks of the range.
// SEL: which selection (BoxSelection, the statistics scan, the caller's).
enum { kSelBox = 0, kSelStats = 1, kSelGiven = 2 };

template <int SEL, bool NU = false>
__global__ void __launch_bounds__(256) select_kernel(BlockRange r, Bc7Params prm, Work w,
                                                     const uint32_t *list,
                                                     const uint32_t *list_count) {
  constexpr int kSlots = SEL == kSelBox ? kMaxBoxCands : kMaxCands;
  __shared__ uint32_t s_px[16][16];
  __shared__ uint32_t s_sel[16];
  __shared__ uint16_t s_cd[16][kSlots];
  __shared__ uint32_t s_cnt[kNumTypes], s_base[kNumTypes];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t base = blockIdx.x * 16;  // position in the list or the range
  const uint32_t total = list ? *list_count : r.num_blocks;
  if (base >= total) return;  // whole workgroup
  auto rel_of = [&](uint32_t i) { return list ? list[i] : i; };
  {
    const int blk = tid >> 4, p = tid & 15;
    const uint32_t i = base + blk;
    if (i < total) {
      const uint32_t rel = rel_of(i);
      const uint32_t b = r.first_block + rel;
      const uint32_t bx = b % r.blocks_wide, by = b / r.blocks_wide;
      s_px[blk][p] = r.img[(size_t)(by * 4 + (p >> 2)) * r.width + bx * 4 + (p & 3)];
    }
  }
  __syncthreads();
  for (int k = 0; k < 4; k++) {
    const int blk = wave * 4 + k;
    if (base + blk >= total) break;  // uniform in the wave
    const uint32_t rel = rel_of(base + blk);
    uint32_t px[16];
#pragma unroll
    for (int i = 0; i < 16; i++) px[i] = s_px[blk][i];
    uint16_t cd[kSlots];
    uint32_t sel = SEL == kSelStats ? stats_select_block(px, lane, prm, cd, prm.stats + rel)
                 : SEL == kSelGiven ? given_select_block(px, prm, prm.selections[rel], cd)
                                    : box_select_block<NU>(px, lane, prm, cd);
    // the learned search's low-confidence blocks (the fallback list) run the
    // full search
    if (list && sel_cls(sel) == 2) sel |= kSelFull;
    const int count = sel_count(sel);
    if (lane == 0) {
      s_sel[blk] = sel;
      w.sel[rel] = sel;
    }
    // slot c of this block's descriptors goes out from lane c
    if (lane < count) w.cdesc[(size_t)rel * kMaxCands + lane] = cd[lane];
    if (lane < kSlots) s_cd[blk][lane] = lane < count ? cd[lane] : (uint16_t)0;
  }
  __syncthreads();
  // enqueue: thread = (block, candidate slot), 16 slots per round
#pragma unroll
  for (int half = 0; half < kSlots / 16; half++) {
    const int blk = tid >> 4, c = (tid & 15) + 16 * half;
    bool valid = false;
    int type = 0;
    uint32_t rel = 0;
    if (base + blk < total && c < sel_count(s_sel[blk])) {
      // mode 6's second shape entry has the same quality-0 fit as the first
      // (its slot follows it): the set-up of the first writes both
      const uint32_t d = s_cd[blk][c];
      valid = !((d & 7) == 6 && ((d >> 12) & 7) == 1);
      rel = rel_of(base + blk);
      type = kTypeOfMode[d & 7];
    }
    wg_enqueue(w, valid, type, cand_id(rel, (uint32_t)c), s_cnt, s_base);
  }
}

// The mode of each written block (the lowest set bit of its first byte; 8
// for the reserved encoding) into its statistics.
__global__ void __launch_bounds__(256) stats_mode_kernel(BlockRange r, fh_bc7_block_stats *stats) {
  const uint32_t rel = blockIdx.x * 256 + threadIdx.x;
  if (rel >= r.num_blocks) return;
  const uint32_t b0 = r.out[(size_t)(r.first_block + rel) * 16];
  stats[rel].mode = b0 ? __ffs(b0) - 1 : 8;
}

// ---------------------------------------------------------------------------
// Kernel 1b (learned candidate pruning).  The ranker gives each block 512
// logits, logit = mode * 64 + slot; these name the real candidates among
// them and the candidate each one stands for.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool logit_real(int j) {
  const int mode = j >> 6, slot = j & 63;
  switch (mode) {
    case 0: return slot < 16;
    case 4: return slot < 8;
    case 5: return slot < 4;
    case 6: return slot == 0;
    default: return true;
  }
}

__device__ __forceinline__ uint32_t cdesc_of_logit(int j) {
  const int mode = j >> 6, slot = j & 63;
  const int shape = kModes[mode].subsets > 1 ? slot : 0;
  const int rot = mode == 4 ? (slot >> 1) : (mode == 5 ? slot : 0);
  const int im = mode == 4 ? (slot & 1) : 0;
  return cdesc_pack(mode, shape, rot, im, 0);
}


// ---------------------------------------------------------------------------
// The learned search's ranking and selection in one kernel: a workgroup of
// four waves takes 64 blocks through the ranker (ranker_tile.h: logits
// bit-identical to the forward pass) and makes each block's decision from
// the fp32 logits staged in LDS, so they never reach HBM.  The output layer
// runs in four phases of 128 logits; the softmax is accumulated online over
// the phases (its total differs from a single sum by rounding only) and the
// top-k list is exact (largest logit first, the lower logit on ties).  Solid and transparent blocks are
// marked as the reference writes them, a confident block (top probability
// at least the threshold) keeps its top-k candidates, and the rest go on the
// fallback list for the full search.
// ---------------------------------------------------------------------------
// Strict order of the top-k list: the larger logit first, the lower logit
// index on ties.
// (Bitwise operators keep the comparisons free of branches.)
__device__ __forceinline__ bool top_before(float va, int ja, float vb, int jb) {
  return (va > vb) | ((va == vb) & (ja < jb));
}

// The same order as one unsigned 64-bit key: the logit's bits made
// monotone (-0 counted as +0, as the float comparisons do) above the
// complemented index, so a larger key comes first.
__device__ __forceinline__ uint64_t top_key(float v, int j) {
  const uint32_t u = __float_as_uint(v + 0.0f);
  const uint32_t o = u ^ ((uint32_t)((int32_t)u >> 31) | 0x80000000u);
  return ((uint64_t)o << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)j);
}
__device__ __forceinline__ float key_value(uint64_t k) {
  const uint32_t o = (uint32_t)(k >> 32);
  return __uint_as_float(o ^ ((uint32_t)((int32_t)~o >> 31) | 0x80000000u));
}
__device__ __forceinline__ int key_index(uint64_t k) {
  return (int)(0xFFFFFFFFu - (uint32_t)k);
}

// Logits per mode that name a real candidate (logit_real).
__device__ __forceinline__ int real_logits(int mode) {
  return mode == 0 ? 16 : (mode == 4 ? 8 : (mode == 5 ? 4 : (mode == 6 ? 1 : 64)));
}

constexpr int kStageStride = 144;  // fp32 logits per staged row: 128 + 16, so that the
                                   // row-quad reads and the wave's writes avoid bank conflicts
static_assert(mlp::kRankRows * kStageStride * 4 <= mlp::kLdsH2, "staging over X and H1");

// Two workgroups per CU (their LDS fits twice): the bound holds the kernel to
// the registers that allows, so one workgroup's MFMA work overlaps the
// other's selection.
__global__ void __launch_bounds__(256, 2) rank_select_kernel(BlockRange r, Bc7Params prm, Work w,
                                                          const uint16_t *w1, const uint16_t *w2,
                                                          const uint16_t *w3,
                                                          const float *params,
                                                          uint32_t *fallback,
                                                          uint32_t *fallback_count) {
  __shared__ __align__(16) uint8_t lds[mlp::kLdsBytes];
  __shared__ uint32_t s_cnt[kNumTypes], s_base[kNumTypes];
  __shared__ uint32_t s_fb, s_fbbase;
  __shared__ uint16_t s_cd[mlp::kRankRows][8];
  __shared__ uint32_t s_count[mlp::kRankRows];
  __shared__ uint32_t s_fbslot[mlp::kRankRows];
  uint16_t *xs = reinterpret_cast<uint16_t *>(lds + mlp::kLdsX);
  uint16_t *h1 = reinterpret_cast<uint16_t *>(lds + mlp::kLdsH1);
  uint16_t *h2 = reinterpret_cast<uint16_t *>(lds + mlp::kLdsH2);
  float *stage = reinterpret_cast<float *>(lds);  // over X and H1 once layer 2 is done
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t n = r.num_blocks;
  const uint32_t tile0 = blockIdx.x * mlp::kRankRows;
  if (tid == 0) s_fb = 0;
  mlp::rank_features(r, n, tile0, tid, xs);
  __syncthreads();
  mlp::rank_hidden<64, mlp::kXs, 2>(xs, w1, params + mlp::kRankOffB1, wave * 64, lane, h1);
  __syncthreads();
  mlp::rank_hidden<256, mlp::kHs, 2>(h1, w2, params + mlp::kRankOffB2, wave * 64, lane, h2);
  __syncthreads();

  // Block srow's decision is made by its four threads (seg 0..3), which hold
  // the same running softmax (m, sum) and top-k list (lv, lj) throughout.
  const int srow = tid >> 2, seg = tid & 3;
  const int kk = prm.topk < 1 ? 1 : (prm.topk > 8 ? 8 : (int)prm.topk);
  const float *b3 = params + mlp::kRankOffB3;
  float m = -INFINITY, sum = 0.0f;
  // the list as keys; an empty entry is (-inf, 2^30)
  const uint64_t kFill = top_key(-INFINITY, 1 << 30);
  uint64_t lk[8];
#pragma unroll
  for (int i = 0; i < 8; i++) lk[i] = kFill;
  // the output layer in four phases of 128 logits (modes 2p and 2p + 1),
  // 32 per wave; each phase is staged in LDS and folded into the row state
#pragma unroll 1
  for (int p = 0; p < 4; p++) {
    mlp::floatx16 acc[2][1] = {};
    const int c0 = 128 * p + 32 * wave;
    mlp::rank_mma<256, mlp::kHs, 1>(h2, w3, c0, lane, acc);
    const float bias = b3[c0 + (lane & 31)];
    if (p > 0) __syncthreads();  // the previous phase's stage is read
#pragma unroll
    for (int rb = 0; rb < 2; rb++)
#pragma unroll
      for (int v = 0; v < 16; v++)
        stage[mlp::acc_row(rb, v, lane) * kStageStride + 32 * wave + (lane & 31)] =
            acc[rb][0][v] + bias;
    __syncthreads();
    // this thread's logits: 128 p + 16 i + 4 seg + e (i < 8, e < 4)
    float z[32];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const float4 q =
          *reinterpret_cast<const float4 *>(stage + srow * kStageStride + 16 * i + 4 * seg);
      z[4 * i] = q.x;
      z[4 * i + 1] = q.y;
      z[4 * i + 2] = q.z;
      z[4 * i + 3] = q.w;
    }
    const int na = ((prm.block_modes >> (2 * p)) & 1) ? real_logits(2 * p) : 0;
    const int nb = ((prm.block_modes >> (2 * p + 1)) & 1) ? real_logits(2 * p + 1) : 0;
    // a logit that names no real candidate of an allowed mode drops out as
    // -inf: never the maximum, exp() = 0 in the sum, never in the top-k
#pragma unroll
    for (int k = 0; k < 32; k++) {
      const int slot = 16 * ((k >> 2) & 3) + 4 * seg + (k & 3);
      z[k] = slot < (k < 16 ? na : nb) ? z[k] : -INFINITY;
    }
    float pm = -INFINITY;
#pragma unroll
    for (int k = 0; k < 32; k++) pm = fmaxf(pm, z[k]);
    pm = fmaxf(pm, __shfl_xor(pm, 1));
    pm = fmaxf(pm, __shfl_xor(pm, 2));
    const float mn = fmaxf(m, pm);
    const float shift = mn > -INFINITY ? mn : 0.0f;
    // the hardware exponential (v_exp_f32 of x log2 e, relative error about
    // 2^-21): the top probability 1 / sum is a threshold test, and its
    // rounding (like that of the summation order) only matters for blocks
    // within about 1e-6 of the threshold
    float ps = 0.0f;
#pragma unroll
    for (int k = 0; k < 32; k++) ps += __expf(z[k] - shift);
    ps += __shfl_xor(ps, 1);
    ps += __shfl_xor(ps, 2);
    if (mn > -INFINITY) {  // the same in the row's four threads
      sum = (m > -INFINITY ? sum * __expf(m - mn) : 0.0f) + ps;
      m = mn;
    }
    // The list changes only if a logit of this phase beats its kk-th entry
    // (an equal one comes later in list order: its index is larger): when
    // no row of the wave has one, as in most waves in the mode-2/3 phase, the
    // merge below would rebuild the same list.
    uint64_t kthk = kFill;
#pragma unroll
    for (int i = 0; i < 8; i++) kthk = i == kk - 1 ? lk[i] : kthk;
    if (!__ballot(pm > key_value(kthk))) continue;
    // this phase's logits as keys (index 128 p + 16 i + 4 seg + e); a masked
    // one gets key 0, below every real key and the empty entry, so it is
    // never taken
    const int jbase = 128 * p + 4 * seg;
    uint64_t zk[32];
#pragma unroll
    for (int k = 0; k < 32; k++)
      zk[k] = z[k] > -INFINITY ? top_key(z[k], jbase + 16 * (k >> 2) + (k & 3)) : 0ull;
    // top-kk of the running list and this phase's logits: the t-th entry is
    // the largest key below entry t - 1's
    uint64_t nk[8];
#pragma unroll
    for (int i = 0; i < 8; i++) nk[i] = kFill;
    uint64_t pk = ~0ull;
#pragma unroll 1
    for (int t = 0; t < kk; t++) {
      uint64_t bk = kFill;
#pragma unroll
      for (int k = 0; k < 32; k++) {
        const bool take = (zk[k] < pk) & (zk[k] > bk);
        bk = take ? zk[k] : bk;
      }
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const bool take = (seg == 0) & (lk[i] < pk) & (lk[i] > bk);
        bk = take ? lk[i] : bk;
      }
#pragma unroll
      for (int off = 1; off < 4; off <<= 1) {
        const uint64_t ok = __shfl_xor(bk, off);
        bk = ok > bk ? ok : bk;
      }
#pragma unroll
      for (int i = 0; i < 8; i++) nk[i] = i == t ? bk : nk[i];
      pk = bk;
    }
#pragma unroll
    for (int i = 0; i < 8; i++) lk[i] = nk[i];
  }
  float lv[8];
  int lj[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    lv[i] = key_value(lk[i]);
    lj[i] = key_index(lk[i]);
  }

  // solid and transparent blocks, from the row's four texel rows
  const uint32_t rel = tile0 + (uint32_t)srow;
  uint4 px = make_uint4(0, 0, 0, 0);
  if (rel < n) {
    const uint32_t b = r.first_block + rel;
    const uint32_t bx = b % r.blocks_wide, by = b / r.blocks_wide;
    px = *reinterpret_cast<const uint4 *>(r.img + (size_t)(by * 4 + seg) * r.width + bx * 4);
  }
  const uint32_t p0 = __shfl(px.x, lane & ~3);
  uint32_t flags = ((px.x != p0 || px.y != p0 || px.z != p0 || px.w != p0) ? 1u : 0u) |
                   (((px.x | px.y | px.z | px.w) >> 24) ? 2u : 0u);
  flags |= __shfl_xor(flags, 1);
  flags |= __shfl_xor(flags, 2);
  uint32_t count = 0, fbslot = 0;
  if (seg == 0 && rel < n) {
    if (!(flags & 1u) || !(flags & 2u)) {
      w.sel[rel] = sel_pack((flags & 1u) ? 1 : 0, 0);  // solid, else transparent
    } else if (!(m > -INFINITY) || 1.0f / sum < prm.threshold) {
      // top probability = exp(0) / sum below the threshold: the full search
      fbslot = atomicAdd(&s_fb, 1u) + 1u;  // 1-based: 0 marks no fallback
    } else {
#pragma unroll
      for (int i = 0; i < 8; i++) {
        if (i < kk && lv[i] > -INFINITY) {
          const uint16_t d = (uint16_t)cdesc_of_logit(lj[i]);
          s_cd[srow][i] = d;
          w.cdesc[(size_t)rel * kMaxCands + i] = d;
          count++;
        }
      }
      w.sel[rel] = sel_pack(2, (int)count);
    }
  }
  if (seg == 0) {
    s_count[srow] = count;
    s_fbslot[srow] = fbslot;
  }
  __syncthreads();
  // fallback list: one global atomic per workgroup
  if (tid == 0 && s_fb) s_fbbase = atomicAdd(fallback_count, s_fb);
  __syncthreads();
  if (tid < mlp::kRankRows && s_fbslot[tid])
    fallback[s_fbbase + s_fbslot[tid] - 1u] = tile0 + (uint32_t)tid;
  // enqueue the confident blocks' candidates: 64 rows x 8 slots
  for (int u = 0; u < 2; u++) {
    const int item = 256 * u + tid;
    const int row = item >> 3, slot = item & 7;
    const bool valid = (uint32_t)slot < s_count[row];
    const int type = valid ? kTypeOfMode[s_cd[row][slot] & 7] : 0;
    wg_enqueue(w, valid, type, cand_id(tile0 + (uint32_t)row, (uint32_t)slot), s_cnt, s_base);
  }
}

// ---------------------------------------------------------------------------
// Kernel 2: candidate set-up, persistent over the type queues in order.
// ---------------------------------------------------------------------------
// Quality-0 fit of one candidate: every subset's endpoints, indices and
// error (CompressionMode::Compress, Compressor.cpp:1300-1371, without the
// annealing), the record the pack kernel reads, and the annealing start
// state.  Returns the mask of subsets that annealing continues (quality > 0).
// MODE >= 0 specialises the fit for that mode at compile time (k.mode must
// equal it); -1 reads the mode from k.
//
// NU: the nonuniform metric with unrotated weights *mu.  The subset errors
// are float bit patterns (a single-colour subset's is the per-pixel error,
// flagged in st.pbc bit 8 + subset, its error being n times it), the alpha
// term a double in st.extra_err / st.pad, and the candidate's total, the
// double sum the reference compares (Compressor.cpp:1300-1371), in *tot_nu.
template <int MODE, bool NU = false>
__device__ __forceinline__ uint32_t fit_candidate(const Cand &k, const uint32_t px[16],
                                                  uint32_t quality, const uint32_t *sctab,
                                                  CandRec &rec, AState &st, uint32_t &total,
                                                  const Weights *mu, double *tot_nu,
                                                  float *km) {
  Fit f;
  const ModeAttr attr = MODE >= 0 ? kModeTab[MODE < 0 ? 0 : MODE] : kModes[k.mode];
  const int nsub = attr.subsets;
  // a specialised mode without rotation or index-mode bits fixes both at
  // zero, so only its own index width's k-means is compiled
  const int rot = (MODE >= 0 && !attr.has_rot) ? 0 : k.rot;
  const int im = (MODE >= 0 && MODE != 4) ? 0 : k.im;
  set_fit_attr(f, attr, MODE >= 0 ? MODE : k.mode, rot, im);
  rec.err = 0;
  rec.px1[0] = rec.px1[1] = rec.px1[2] = 0;
  rec.px2[0] = rec.px2[1] = rec.px2[2] = 0;
  rec.aidx_lo = rec.aidx_hi = 0;
  st.extra_err = 0;
  st.pad = 0;
  uint32_t active = 0, pbcs = 0, scf = 0;
  total = 0;
  double tnu = 0.0;
  uint64_t cidx = 0;
  if (attr.has_rot) {
    uint64_t aidx;
    uint32_t a1q, a2q;
    Weights mr;
    double aerr_nu = 0.0;
    if constexpr (NU) mr = rotate_weights(*mu, rot);
    const uint32_t aerr = alpha_fit<NU>(f, px, aidx, a1q, a2q, NU ? mr.w[3] : 1.0f, &aerr_nu);
    // the RGB fit sees the stale statistics of the unrotated cluster
    float avg[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < 16; i++)
#pragma unroll
      for (int kk = 0; kk < 4; kk++) avg[kk] += (float)chan(px[i], kk);
#pragma unroll
    for (int kk = 0; kk < 4; kk++) avg[kk] = avg[kk] / 16.0f;
    SubStart s0;
    subset_setup<NU>(f, quality, sctab, px, 0xFFFFu, rot + 1, 16, avg, false, s0, mu, &mr, km);
    rec.aidx_lo = (uint32_t)aidx;
    rec.aidx_hi = (uint32_t)(aidx >> 32);
    // Pack writes the RGB of the fit and the fitted alpha, no p-bit
    rec.px1[0] = (to_pixel_int(s0.p1, f.qmask, -1) & 0x00FFFFFFu) | (a1q << 24);
    rec.px2[0] = (to_pixel_int(s0.p2, f.qmask, -1) & 0x00FFFFFFu) | (a2q << 24);
    st.p1[0] = s0.p1;
    st.p2[0] = s0.p2;
    st.err[0] = s0.err;
    st.extra_err = aerr;
    if constexpr (NU) {
      // rgbError + alphaError (Compressor.cpp:918)
      const double rgb = s0.sc ? 16.0 * (double)__uint_as_float(s0.err)
                               : (double)__uint_as_float(s0.err);
      tnu = rgb + aerr_nu;
      const uint64_t ab = (uint64_t)__double_as_longlong(aerr_nu);
      st.extra_err = (uint32_t)ab;
      st.pad = (uint32_t)(ab >> 32);
      scf = s0.sc ? 1u : 0u;
    }
    cidx = s0.idx;
    total = s0.err + aerr;
    active = s0.anneal ? 1u : 0u;
  } else {
#pragma unroll 1
    for (int s = 0; s < nsub; s++) {
      const uint32_t smask = subset_mask(k.shape, nsub, s);
      const int n = __popc(smask);
      float avg[4] = {0.0f, 0.0f, 0.0f, 0.0f};
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
        for (int kk = 0; kk < 4; kk++) avg[kk] += (float)chan(px[i], kk);
      }
      // integer sums over n in 1..16: the double product is the correctly
      // rounded quotient (as for the k-means means)
#pragma unroll
      for (int kk = 0; kk < 4; kk++) avg[kk] = (float)((double)avg[kk] * kInvCount[n]);
      SubStart ss;
      subset_setup<NU>(f, quality, sctab, px, smask, 0, n, avg, same, ss, mu, mu, km);
      if constexpr (NU) {
        tnu += ss.sc ? (double)n * (double)__uint_as_float(ss.err)
                     : (double)__uint_as_float(ss.err);
        scf |= (ss.sc ? 1u : 0u) << s;
      }
      const int pb0 = f.pbit_type == PB_NONE ? -1 : pbc_bit(f.pbit_type, ss.pbc, 0);
      const int pb1 = f.pbit_type == PB_NONE ? -1 : pbc_bit(f.pbit_type, ss.pbc, 1);
      const uint32_t e1 = to_pixel_int(ss.p1, f.qmask, pb0);
      const uint32_t e2 = to_pixel_int(ss.p2, f.qmask, pb1);
      if (s == 0) {
        rec.px1[0] = e1; rec.px2[0] = e2;
        st.p1[0] = ss.p1; st.p2[0] = ss.p2; st.err[0] = ss.err;
      } else if (s == 1) {
        rec.px1[1] = e1; rec.px2[1] = e2;
        st.p1[1] = ss.p1; st.p2[1] = ss.p2; st.err[1] = ss.err;
      } else {
        rec.px1[2] = e1; rec.px2[2] = e2;
        st.p1[2] = ss.p1; st.p2[2] = ss.p2; st.err[2] = ss.err;
      }
      pbcs |= (uint32_t)ss.pbc << (2 * s);
      cidx |= ss.idx;
      total += ss.err;
      if (ss.anneal) active |= 1u << s;
    }
  }
  rec.cidx_lo = (uint32_t)cidx;
  rec.cidx_hi = (uint32_t)(cidx >> 32);
  rec.meta = pbcs | ((uint32_t)rot << 8) | ((uint32_t)im << 10);
  rec.err = total;  // the quality-0 result; annealing overwrites active ones
  st.pbc = pbcs;
  if constexpr (NU) {
    st.pbc |= scf << 8;
    *tot_nu = tnu;
  }
  st.idx_lo = (uint32_t)cidx;
  st.idx_hi = (uint32_t)(cidx >> 32);
  return active;
}

// One launch per mode (its type queue), specialised at compile time so that
// each mode gets its own register allocation.
constexpr int kPxStride = 17;  // odd stride: lanes reading texel i hit distinct banks

// Index entries of the widest k-means a mode's set-up runs (the colour
// indices; mode 4 with index mode 1 has 3-bit colour indices).
constexpr int kSetupBuckets(int mode) {
  return mode == 6 ? 16 : ((mode == 0 || mode == 1 || mode == 4) ? 8 : 4);
}

template <int MODE, bool NU = false>
__global__ void __launch_bounds__(64) setup_kernel(BlockRange r, Bc7Params prm, Work w,
                                                   const uint32_t *sctab) {
  const int lane = threadIdx.x;
  __shared__ uint32_t s_px[64 * kPxStride];
  // the k-means sums and counts, per bucket and channel a row of 64 lanes
  __shared__ float s_km[kSetupBuckets(MODE) * 5 * 64];
  {
    constexpr int kTypeOf[8] = {0, 2, 1, 3, 6, 7, 5, 4};
    const int t = kTypeOf[MODE];
    const uint32_t count = *(volatile uint32_t *)&w.counters[t];
    for (;;) {
      uint32_t b0 = 0;
      if (lane == 0) b0 = atomicAdd(&w.counters[8 + t], 64u);
      b0 = __shfl(b0, 0);
      if (b0 >= count) break;
      const uint32_t item = b0 + lane;
      if (item >= count) continue;
      const uint32_t id = w.queue[(size_t)t * w.qcap + item];
      const uint32_t rel = id_block(id);
      const int c = (int)id_slot(id);
      Cand k;
      cand_from_desc(w.cdesc[(size_t)rel * kMaxCands + c], k);
      // the lane's texels live in LDS: the fit reads them many times, and
      // keeping them (and values derived from them) out of registers lets
      // the loops below stay rolled
      {
        uint32_t v[16];
        load_block(r, rel, v);
#pragma unroll
        for (int i = 0; i < 16; i++) s_px[lane * kPxStride + i] = v[i];
      }
      const uint32_t *px = s_px + lane * kPxStride;
      CandRec rec;
      AState st;
      uint32_t total = 0;
      double tnu = 0.0;
      Weights mu;
      if constexpr (NU) mu = params_weights(prm);
      const uint32_t active =
          fit_candidate<MODE, NU>(k, px, prm.quality, sctab, rec, st, total, &mu, &tnu,
                                  s_km + lane);
      st.id = id;
      w.rec[(size_t)rel * kMaxCands + c] = rec;
      if constexpr (NU) w.nu_total[(size_t)rel * kMaxCands + c] = tnu;
      // mode 6, shape entry 0 followed by entry 1 (opaque blocks with two
      // shapes): the second slot gets the same fit and its own annealing state
      bool twin = false;
      if (MODE == 6 && c + 1 < kMaxCands && c + 1 < sel_count(w.sel[rel])) {
        const uint32_t d1 = w.cdesc[(size_t)rel * kMaxCands + c + 1];
        twin = (d1 & 7) == 6 && ((d1 >> 12) & 7) == 1;
      }
      if (twin) {
        w.rec[(size_t)rel * kMaxCands + c + 1] = rec;
        if constexpr (NU) w.nu_total[(size_t)rel * kMaxCands + c + 1] = tnu;
      }
      if (active) {
        st.info = (uint32_t)k.mode | ((uint32_t)k.rot << 3) | ((uint32_t)k.im << 5) |
                  ((uint32_t)k.nsub << 6) | ((uint32_t)k.shape << 8) | (active << 14) |
                  ((uint32_t)k.entry << 17);
        const uint32_t slot = atomicAdd(&w.counters[16], twin ? 2u : 1u);
        w.astate[slot] = st;
        if (w.aid) w.aid[slot] = id;
        if (twin) {
          st.id = id + 1;
          st.info |= 1u << 17;
          w.astate[slot + 1] = st;
          if (w.aid) w.aid[slot + 1] = id + 1;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Kernel 3: simulated annealing (OptimizeEndpointsForCluster,
// Compressor.cpp:583-624), persistent.  A lane owns one candidate and
// advances every still-annealing subset of it by one step per iteration; the
// 16 pixels are scored in one pass, each against its own subset's
// endpoints.  Lanes that finish take the next candidate from a wave-local
// window of the global list, so the wave stays full.
// ---------------------------------------------------------------------------
struct Lane {
  bool has;
  uint32_t id;
  Fit f;
  int nsub, entry, shape;
  uint32_t sub2;       // subset of pixel i in bits 2i..2i+1
  int ptmode;
  uint32_t *px;        // the block's 16 texels: this lane's row of LDS
  uint32_t *pt;        // the points the RGB fit sees (rotation applied), same row
  uint32_t active;
  uint32_t p1[3], p2[3], bp1[3], bp2[3], cur[3], best[3];
  int cpbc[3], bpbc[3], energy[3];
  uint32_t key[3], ctr[3];
  uint64_t idx;        // indices of all subsets (best so far)
  uint32_t extra;
  // nonuniform metric: the candidate's weights (its rotation applied), the
  // single-colour subsets (their error is n times the stored one) and the
  // alpha term
  Weights m;
  uint32_t scf;
  double extra_nu;
};

template <bool NU = false>
__device__ void lane_load(Lane &L, const BlockRange &r, const Bc7Params &prm, const AState &st) {
  L.has = true;
  L.id = st.id;
  const int mode = st.info & 7, rot = (st.info >> 3) & 3, im = (st.info >> 5) & 1;
  L.nsub = (st.info >> 6) & 3;
  L.shape = (st.info >> 8) & 63;
  L.active = (st.info >> 14) & 7;
  L.entry = (st.info >> 17) & 7;
  set_fit(L.f, mode, rot, im);
  L.ptmode = kModes[mode].has_rot ? rot + 1 : 0;
  L.sub2 = 0;
  for (int s = 1; s < L.nsub; s++) {
    const uint32_t m = subset_mask(L.shape, L.nsub, s);
#pragma unroll
    for (int i = 0; i < 16; i++)
      if ((m >> i) & 1) L.sub2 |= (uint32_t)s << (2 * i);
  }
  {
    uint32_t v[16];
    load_block(r, id_block(st.id), v);
#pragma unroll
    for (int i = 0; i < 16; i++) {
      L.px[i] = v[i];
      L.pt[i] = point_of(v[i], L.ptmode);
    }
  }
  const uint64_t bid = prm.block_id_base + r.first_block + id_block(st.id);
#pragma unroll
  for (int s = 0; s < 3; s++) {
    L.p1[s] = L.bp1[s] = st.p1[s];
    L.p2[s] = L.bp2[s] = st.p2[s];
    L.cur[s] = L.best[s] = st.err[s];
    L.cpbc[s] = L.bpbc[s] = (st.pbc >> (2 * s)) & 3;
    L.energy[s] = 0;
    L.ctr[s] = 0;
    L.key[s] = fit_key(prm.seed, bid, cand_tag(mode, L.entry, s, rot, im));
  }
  L.idx = (uint64_t)st.idx_lo | ((uint64_t)st.idx_hi << 32);
  L.extra = st.extra_err;
  if constexpr (NU) {
    L.m = rotate_weights(params_weights(prm), rot);
    L.scf = (st.pbc >> 8) & 7u;
    L.extra_nu = __longlong_as_double((long long)((uint64_t)st.extra_err |
                                                  ((uint64_t)st.pad << 32)));
  }
}

// The 16-pixel pass for up to NSUB subsets: errors per subset and the
// indices.  Every pixel is scored; errors of subsets that have stopped
// annealing are not used.
template <int NSUB>
__device__ __forceinline__ void score_pixels(const Lane &L, const SubConst &sc0,
                                             const SubConst &sc1, const SubConst &sc2, int nb,
                                             float kf, WeightFn wf, uint32_t nerr[3],
                                             uint64_t &nidx) {
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const uint32_t s = NSUB == 1 ? 0u : (L.sub2 >> (2 * i)) & 3u;
    const SubConst c = NSUB == 1 ? sc0
                                 : (NSUB == 2 ? pick_sub(s != 0, sc1, sc0)
                                              : pick_sub(s == 0, sc0, pick_sub(s == 1, sc1, sc2)));
    const uint32_t pix = L.px[i];
    const uint32_t pt = L.pt[i];
    int best;
    const uint32_t me = pixel_err(pix, pt, c, nb, kf, wf, &best);
    if (NSUB == 1) {
      nerr[0] += me;
    } else {
      nerr[0] += s == 0 ? me : 0u;
      nerr[1] += s == 1 ? me : 0u;
      if (NSUB == 3) nerr[2] += s == 2 ? me : 0u;
    }
    nidx |= (uint64_t)best << (4 * i);
  }
}

// The pixel pass under the nonuniform metric: float errors, each subset's
// summed in pixel order (adding 0 to the other subsets' sums leaves them
// unchanged).
template <int NSUB>
__device__ __forceinline__ void score_pixels_nu(const Lane &L, const SubConst &sc0,
                                                const SubConst &sc1, const SubConst &sc2, int nb,
                                                WeightFn wf, float nerr[3], uint64_t &nidx) {
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const uint32_t s = NSUB == 1 ? 0u : (L.sub2 >> (2 * i)) & 3u;
    const SubConst c = NSUB == 1 ? sc0
                                 : (NSUB == 2 ? pick_sub(s != 0, sc1, sc0)
                                              : pick_sub(s == 0, sc0, pick_sub(s == 1, sc1, sc2)));
    int best;
    const float me = pixel_err_nu(L.px[i], L.pt[i], c, nb, wf, L.m, &best);
    nerr[0] += s == 0 ? me : 0.0f;
    if (NSUB > 1) nerr[1] += s == 1 ? me : 0.0f;
    if (NSUB == 3) nerr[2] += s == 2 ? me : 0.0f;
    nidx |= (uint64_t)best << (4 * i);
  }
}

// The same pass split over a group of G lanes (G = 2, 4, 8 or 16, aligned in
// the wave) that hold the same candidate: member m scores pixels m, m + G,
// ...; the sums and index bits are combined over the group.  The per-subset
// errors are integers, so the split does not change them.
template <int NSUB>
__device__ __forceinline__ void score_pixels_group(const Lane &L, const SubConst &sc0,
                                                   const SubConst &sc1, const SubConst &sc2,
                                                   int nb, float kf, WeightFn wf, uint32_t nerr[3],
                                                   uint64_t &nidx, int G, int m) {
  for (int i = m; i < 16; i += G) {
    const uint32_t s = NSUB == 1 ? 0u : (L.sub2 >> (2 * i)) & 3u;
    const SubConst c = NSUB == 1 ? sc0
                                 : (NSUB == 2 ? pick_sub(s != 0, sc1, sc0)
                                              : pick_sub(s == 0, sc0, pick_sub(s == 1, sc1, sc2)));
    int best;
    const uint32_t me = pixel_err(L.px[i], L.pt[i], c, nb, kf, wf, &best);
    nerr[0] += s == 0 ? me : 0u;
    if (NSUB > 1) nerr[1] += s == 1 ? me : 0u;
    if (NSUB == 3) nerr[2] += s == 2 ? me : 0u;
    nidx |= (uint64_t)best << (4 * i);
  }
  uint32_t lo = (uint32_t)nidx, hi = (uint32_t)(nidx >> 32);
  for (int off = 1; off < G; off <<= 1) {
    nerr[0] += __shfl_xor(nerr[0], off);
    if (NSUB > 1) nerr[1] += __shfl_xor(nerr[1], off);
    if (NSUB == 3) nerr[2] += __shfl_xor(nerr[2], off);
    lo |= __shfl_xor(lo, off);
    hi |= __shfl_xor(hi, off);
  }
  nidx = (uint64_t)lo | ((uint64_t)hi << 32);
}

// One annealing step of every active subset.  Returns true when the
// candidate is finished.  With GROUPED, the lanes of an aligned group of G
// hold the same candidate and share the pixel pass (member m); every other
// part of the step runs identically in each of them.
template <bool GROUPED = false, bool NU = false>
__device__ bool lane_step(Lane &L, uint32_t q, int G = 1, int m = 0) {
  const Fit &f = L.f;
  const int nb = 1 << f.nbits_idx;
  uint32_t np1[3], np2[3], e1[3], e2[3];
  int npbc[3];
#pragma unroll
  for (int s = 0; s < 3; s++) {
    np1[s] = np2[s] = e1[s] = e2[s] = 0;
    npbc[s] = 0;
    if (!((L.active >> s) & 1)) continue;
    Rng rng;
    rng.key = L.key[s];
    rng.ctr = L.ctr[s];
    pick_neighbor(f, rng, L.p1[s], L.p2[s], L.cpbc[s], np1[s], np2[s], npbc[s], L.bp1[s],
                  L.bp2[s], L.bpbc[s]);
    L.ctr[s] = rng.ctr;
    e1[s] = to_pixel_tab(np1[s], f.qmask, pbc_bit(f.pbit_type, npbc[s], 0));
    e2[s] = to_pixel_tab(np2[s], f.qmask, pbc_bit(f.pbit_type, npbc[s], 1));
  }
  const SubConst sc0 = sub_const(e1[0], e2[0]);
  const SubConst sc1 = sub_const(e1[1], e2[1]);
  const SubConst sc2 = sub_const(e1[2], e2[2]);
  uint32_t nerr[3] = {0, 0, 0};
  uint64_t nidx = 0;
  const float kf = (float)(nb - 1);
  const WeightFn wf = weight_fn(nb);
  // the widest subset count among the wave's running lanes picks the path
  // (a wave-uniform branch): the list is ordered by mode, so most waves hold
  // candidates of one subset count
  if constexpr (NU) {
    float nerf[3] = {0.0f, 0.0f, 0.0f};
    if (__ballot(L.nsub == 3)) score_pixels_nu<3>(L, sc0, sc1, sc2, nb, wf, nerf, nidx);
    else if (__ballot(L.nsub == 2)) score_pixels_nu<2>(L, sc0, sc1, sc2, nb, wf, nerf, nidx);
    else score_pixels_nu<1>(L, sc0, sc1, sc2, nb, wf, nerf, nidx);
#pragma unroll
    for (int s = 0; s < 3; s++) nerr[s] = __float_as_uint(nerf[s]);
  } else if (GROUPED) {
    if (__ballot(L.nsub == 3)) score_pixels_group<3>(L, sc0, sc1, sc2, nb, kf, wf, nerr, nidx, G, m);
    else if (__ballot(L.nsub == 2)) score_pixels_group<2>(L, sc0, sc1, sc2, nb, kf, wf, nerr, nidx, G, m);
    else score_pixels_group<1>(L, sc0, sc1, sc2, nb, kf, wf, nerr, nidx, G, m);
  } else if (__ballot(L.nsub == 3)) {
    score_pixels<3>(L, sc0, sc1, sc2, nb, kf, wf, nerr, nidx);
  } else if (__ballot(L.nsub == 2)) {
    score_pixels<2>(L, sc0, sc1, sc2, nb, kf, wf, nerr, nidx);
  } else {
    score_pixels<1>(L, sc0, sc1, sc2, nb, kf, wf, nerr, nidx);
  }
#pragma unroll
  for (int s = 0; s < 3; s++) {
    if (!((L.active >> s) & 1)) continue;
    const uint32_t err = nerr[s];
    bool accept = err < L.cur[s];
    if (!accept) {
      // AcceptNewEndpointError (Compressor.cpp:524-536): accept when
      // frand() < exp(0.1 (cur - err) / temp).  Here err >= cur, so the
      // probability is at most 1 and a draw in [1, 3) never accepts.
      Rng rng;
      rng.key = L.key[s];
      rng.ctr = L.ctr[s];
      const float rr = rng.frand();
      L.ctr[s] = rng.ctr;
      if (rr < 1.0f) {
        // A single-precision estimate of the probability decides unless rr
        // lies within 1e-4 (relative) of it, which covers the estimate's
        // error; then the double expression of the reference decides.  The
        // uniform metric's errors are integers below 2^24, so their float
        // difference is exact; the nonuniform metric's are floats, whose
        // difference has a relative error of at most 2^-23 where it is not
        // exact, well inside the margin.
        const float cf = NU ? __uint_as_float(L.cur[s]) : (float)L.cur[s];
        const float ef = NU ? __uint_as_float(err) : (float)err;
        const float diff = cf - ef;
        const float pf =
            __expf(__fdividef(0.1f * diff * (float)((int)q - 1), (float)L.energy[s]));
        if (rr != 0.0f && rr < pf * 0.9999f) {
          accept = true;
        } else if (!(rr != 0.0f && rr > pf * 1.0001f)) {
          const float temp = (float)L.energy[s] / (float)((int)q - 1);
          const double cd = NU ? (double)cf : (double)L.cur[s];
          const double ed = NU ? (double)ef : (double)err;
          const double pr = exp(((double)0.1f * (cd - ed)) / (double)temp);
          accept = (double)rr < pr;
        }
      }
    }
    if (accept) {
      L.cur[s] = err;
      L.p1[s] = np1[s];
      L.p2[s] = np2[s];
      L.cpbc[s] = npbc[s];
    }
    if (err < L.best[s]) {
      // indices of this subset's pixels
      uint64_t m = 0;
#pragma unroll
      for (int i = 0; i < 16; i++)
        if (((L.sub2 >> (2 * i)) & 3) == (uint32_t)s) m |= (uint64_t)0xF << (4 * i);
      L.idx = (L.idx & ~m) | (nidx & m);
      L.bp1[s] = np1[s];
      L.bp2[s] = np2[s];
      L.bpbc[s] = npbc[s];
      L.best[s] = err;
      L.energy[s] = 0;
    }
    L.energy[s]++;
    if (!(L.best[s] > 0 && L.energy[s] < (int)q)) L.active &= ~(1u << s);
  }
  return L.active == 0;
}

template <bool NU = false>
__device__ void lane_finish(Lane &L, Work &w) {
  CandRec *rc = &w.rec[L.id];
  const Fit &f = L.f;
  uint32_t total = L.extra;
  // nonuniform metric: the reference's double total, subsets in order, then
  // the alpha term of a rotation mode (Compressor.cpp:918, 1300-1371)
  double tnu = 0.0;
  uint32_t pbcs = rc->meta & 0x3Fu;
  for (int s = 0; s < L.nsub; s++) {
    const uint32_t bp1 = s == 0 ? L.bp1[0] : (s == 1 ? L.bp1[1] : L.bp1[2]);
    const uint32_t bp2 = s == 0 ? L.bp2[0] : (s == 1 ? L.bp2[1] : L.bp2[2]);
    const int pb = s == 0 ? L.bpbc[0] : (s == 1 ? L.bpbc[1] : L.bpbc[2]);
    const uint32_t be = s == 0 ? L.best[0] : (s == 1 ? L.best[1] : L.best[2]);
    total += be;
    if constexpr (NU) {
      double e = (double)__uint_as_float(be);
      if ((L.scf >> s) & 1) {
        int n = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) n += ((L.sub2 >> (2 * i)) & 3) == (uint32_t)s ? 1 : 0;
        e = (double)n * e;
      }
      tnu += e;
    }
    const int pb0 = f.pbit_type == PB_NONE ? -1 : pbc_bit(f.pbit_type, pb, 0);
    const int pb1 = f.pbit_type == PB_NONE ? -1 : pbc_bit(f.pbit_type, pb, 1);
    uint32_t e1 = to_pixel_int(bp1, f.qmask, pb0);
    uint32_t e2 = to_pixel_int(bp2, f.qmask, pb1);
    if (kModes[f.mode].has_rot) {
      e1 = (e1 & 0x00FFFFFFu) | (rc->px1[0] & 0xFF000000u);
      e2 = (e2 & 0x00FFFFFFu) | (rc->px2[0] & 0xFF000000u);
    }
    rc->px1[s] = e1;
    rc->px2[s] = e2;
    pbcs = (pbcs & ~(3u << (2 * s))) | ((uint32_t)pb << (2 * s));
  }
  rc->cidx_lo = (uint32_t)L.idx;
  rc->cidx_hi = (uint32_t)(L.idx >> 32);
  rc->meta = (rc->meta & ~0x3Fu) | pbcs | kMetaAnnealed;
  rc->err = total;
  if constexpr (NU) w.nu_total[L.id] = tnu + L.extra_nu;
}

// A wave with at most this many candidates left after the list is used up
// finishes them in lane groups.
constexpr int kTailLanes = 16;
// Lanes per candidate for the rotation modes (4 and 5), whose annealing
// chains are the longest (hundreds to over a thousand steps against tens for
// the other modes): run one lane each, they would set the launch's length.
// Pairs split the pixel pass in two; every member repeats the rest of the
// step, so larger groups shorten a step less than they cost (measured: 4
// and 8 lanes slower overall).
constexpr int kRotGroup = 2;
// The wave mask of each group's first lane (lanes 0, g, 2g, ...).
constexpr uint64_t group_leaders(int g, int lane = 0) {
  return lane >= 64 ? 0ull : (1ull << lane) | group_leaders(g, lane + g);
}
static_assert(group_leaders(2) == 0x5555555555555555ull && group_leaders(4) == 0x1111111111111111ull,
              "group leaders");

// UNROLL annealing steps run between refills of finished lanes.
// WAVES: the minimum waves per SIMD the register allocation must allow.  3
// (168 VGPRs, a small spill) suits both searches since the annealing step's
// quantisation became table loads (before, the learned search ran alike at 2
// and 3 waves per SIMD: 7.66 and 7.70 ms; now 6.08 and 5.78 ms).
// NU: the nonuniform metric (float errors; the tail below, which splits a
// candidate's pixel sums over lanes, would change their order and is not
// used: the lanes finish their own candidates).
// COUNT: also count the work into prm.work_counts: [0] the QuantizedError
// point evaluations the reference's annealing makes (each step of a subset
// scores that subset's points, Compressor.cpp:583-624), [1] the pixel scores
// this kernel makes (16 per step of a candidate, whichever subsets are
// still annealing).
template <int UNROLL, int WAVES, bool NU = false, bool COUNT = false>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WAVES)))
anneal_kernel(BlockRange r, Bc7Params prm, Work w) {
  const int lane = threadIdx.x;
  unsigned long long n_ref = 0, n_gpu = 0;  // COUNT
  auto count_step = [&](const Lane &x) {
    // points of the subsets still annealing: the subset codes' low and high
    // bits give the sizes of subsets 1 and 2
    const uint32_t lo = x.sub2 & 0x55555555u, hi = (x.sub2 >> 1) & 0x55555555u;
    const uint32_t n1 = (uint32_t)__popc(lo & ~hi), n2 = (uint32_t)__popc(hi & ~lo);
    const uint32_t n0 = 16u - n1 - n2;
    n_ref += ((x.active & 1u) ? n0 : 0u) + ((x.active & 2u) ? n1 : 0u) +
             ((x.active & 4u) ? n2 : 0u);
    n_gpu += 16u;
  };
  auto flush_counts = [&]() {
    if constexpr (COUNT) {
      for (int off = 32; off >= 1; off >>= 1) {
        n_ref += __shfl_xor(n_ref, off);
        n_gpu += __shfl_xor(n_gpu, off);
      }
      if (lane == 0) {
        atomicAdd(&prm.work_counts[0], n_ref);
        atomicAdd(&prm.work_counts[1], n_gpu);
      }
    }
  };
  const uint64_t lt = (1ull << lane) - 1ull;
  const uint32_t listed = *(volatile uint32_t *)&w.counters[w.alist ? 18 : 16];
  // per lane 16 texels then 16 points, odd row stride: conflict-free
  __shared__ uint32_t s_px[64 * 33];
  // Learned search: the list is in state order, and the set-up kernels
  // append the rotation modes' states last (counters[19] holds the first of
  // them): the list's rotation states are its suffix from rot_begin.  (The
  // full search's list is mostly rotation states, whose grouped steps would
  // cost more than its short tail: it runs them one lane each.)
  uint32_t rot_begin = listed;
  if (!NU && w.alist) {
    const uint32_t first_rot = *(volatile uint32_t *)&w.counters[19];
    uint32_t lo = 0, hi = listed;  // first list entry >= first_rot
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (w.alist[mid] < first_rot) lo = mid + 1;
      else hi = mid;
    }
    rot_begin = lo;
  }
  // Phase A, in the first eighth of the launch (a quarter of the waves that
  // fit at once): the rotation states in groups of kRotGroup lanes that
  // share each step's pixel pass (every member holds the same state and
  // texels), fetched from the end of the suffix.  A group that finishes
  // takes the next; once the suffix is used up the wave's groups finish
  // theirs and the wave joins phase B.  The other waves start on phase B at
  // once, so the longest chains run from the start, with shorter steps,
  // while the rest of the list proceeds beside them.
  if (!NU && rot_begin < listed && blockIdx.x < (gridDim.x + 7) / 8) {
    const uint32_t nrot = listed - rot_begin;
    const int m = lane % kRotGroup;
    constexpr uint64_t leaders = group_leaders(kRotGroup);
    Lane T;
    T.px = s_px + lane * 33;
    T.pt = T.px + 16;
    T.has = false;
    uint32_t abase = 0, aend = 0;
    bool aexh = false;
    for (;;) {
      for (;;) {
        const uint64_t need = __ballot(!T.has) & leaders;
        if (!need || aexh) break;
        if (abase >= aend) {
          uint32_t b0 = 0;
          if (lane == 0) b0 = atomicAdd(&w.counters[21], 16u);
          b0 = __shfl(b0, 0);
          if (b0 >= nrot) {
            aexh = true;
            break;
          }
          abase = b0;
          aend = min(b0 + 16u, nrot);
        }
        const uint32_t avail = aend - abase;
        // rank of this lane's group among the groups that need a state
        const uint64_t below = (1ull << (lane & ~(kRotGroup - 1))) - 1ull;
        const uint32_t rank = (uint32_t)__popcll(need & below);
        const bool wants = ((need >> (lane & ~(kRotGroup - 1))) & 1) != 0;
        if (wants && !T.has && rank < avail) {
          const uint32_t item = listed - 1u - (abase + rank);
          lane_load<NU>(T, r, prm, w.astate[w.alist ? w.alist[item] : item]);
        }
        abase += min(avail, (uint32_t)__popcll(need));
      }
      if (!__ballot(T.has)) break;
      for (int u = 0; u < UNROLL; u++) {
        if (T.has) {
          if constexpr (COUNT) {
            if (m == 0) count_step(T);
          }
          if (lane_step<true>(T, prm.quality, kRotGroup, m)) {
            if (m == 0) lane_finish(T, w);
            T.has = false;
          }
        }
      }
    }
  }
  // Phase B: the other states, one lane each, from the end of the prefix
  const uint32_t total = rot_begin;
  Lane L;
  L.px = s_px + lane * 33;
  L.pt = L.px + 16;
  L.has = false;
  uint32_t wbase = 0, wend = 0;  // wave-local window of the list
  bool exhausted = false;
  for (;;) {
    // the list is used up and few candidates remain: the tail below
    if (!NU && exhausted && __popcll(__ballot(L.has)) <= kTailLanes) break;
    // refill idle lanes from the window, fetching a new window when needed
    for (;;) {
      const uint64_t need = __ballot(!L.has);
      if (!need || exhausted) break;
      if (wbase >= wend) {
        uint32_t b0 = 0;
        if (lane == 0) b0 = atomicAdd(&w.counters[17], 64u);
        b0 = __shfl(b0, 0);
        if (b0 >= total) {
          exhausted = true;
          break;
        }
        wbase = b0;
        wend = min(b0 + 64u, total);
      }
      const uint32_t avail = wend - wbase;
      const uint32_t rank = (uint32_t)__popcll(need & lt);
      if (!L.has && rank < avail) {
        // the list runs from its end: the set-up kernels append modes 0, 2,
        // 3, 7, 6, 1, 4, 5.  The rotation modes' candidates anneal the
        // longest, so the full search starts with them; then come the
        // numerous mode-1 candidates, and the short mode-7 chains are among
        // the last loaded, which keeps the tail of the launch short
        const uint32_t item = total - 1u - (wbase + rank);
        lane_load<NU>(L, r, prm, w.astate[w.alist ? w.alist[item] : item]);
      }
      wbase += min(avail, (uint32_t)__popcll(need));
    }
    if (!__ballot(L.has)) break;
    for (int u = 0; u < UNROLL; u++) {
      if (L.has) {
        if constexpr (COUNT) count_step(L);
        if (lane_step<false, NU>(L, prm.quality)) {
          lane_finish<NU>(L, w);
          L.has = false;
        }
      }
    }
  }
  if constexpr (NU) {
    flush_counts();
    return;
  }
  // Tail: the wave's last candidates (at most kTailLanes) each get a group
  // of G lanes that share the pixel pass of every step, so the longest
  // anneals, which set the launch's length, run up to G times fewer
  // instructions per step.  Their texels stay in the LDS rows of the lanes
  // that loaded them.
  const uint64_t act = __ballot(L.has);
  const int na = __popcll(act);
  if (na == 0) {
    flush_counts();
    return;
  }
  __shared__ Lane s_tail[kTailLanes];
  __shared__ uint32_t s_row[kTailLanes];
  if (L.has) {
    const int k = __popcll(act & lt);
    s_tail[k] = L;
    s_row[k] = (uint32_t)lane;
  }
  __syncthreads();
  const int G = na <= 4 ? 16 : (na <= 8 ? 8 : 4);
  const int g = lane / G, m = lane % G;
  Lane T;
  T.has = false;
  if (g < na) {
    T = s_tail[g];
    T.px = s_px + s_row[g] * 33;
    T.pt = T.px + 16;
  }
  while (__ballot(T.has)) {
    for (int u = 0; u < UNROLL; u++) {
      if (T.has) {
        if constexpr (COUNT) {
          if (m == 0) count_step(T);  // the group scores the 16 pixels once
        }
        if (lane_step<true>(T, prm.quality, G, m)) {
          if (m == 0) lane_finish(T, w);
          T.has = false;
        }
      }
    }
  }
  flush_counts();
}

// ---------------------------------------------------------------------------
// Kernel 4: choose the mode and write the block, one lane per block.
// ---------------------------------------------------------------------------

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

__device__ __forceinline__ uint32_t invert_idx(uint32_t &lo, uint32_t &hi, uint32_t mask16,
                                               uint32_t nv) {
  uint32_t nl = 0, nh = 0;
  for (int i = 0; i < 16; i++) {
    uint32_t v = idx_at(lo, hi, i);
    if ((mask16 >> i) & 1) v = nv - v;
    if (i < 8) nl |= v << (4 * i);
    else nh |= v << (4 * (i - 8));
  }
  lo = nl;
  hi = nh;
  return 0;
}

// CompressionMode::Pack (Compressor.cpp:1096-1298).  The p-bits are written
// unswapped after an anchor fix-up swap, as the reference does
// (:1149-1170 vs :1234-1241).
__device__ __forceinline__ void pack_mode(int mode, int shape, const CandRec &rec, BitWriter &w) {
  const ModeAttr a = kModes[mode];
  const int ns = a.subsets;
  const int rot = (rec.meta >> 8) & 3;
  const int im = (rec.meta >> 10) & 1;
  w.put(1u << mode, mode + 1);
  w.put((uint32_t)shape, a.part_bits);
  if (a.has_rot) w.put((uint32_t)rot, 2);
  if (a.has_idx_mode) w.put((uint32_t)im, 1);
  uint32_t px1[3] = {rec.px1[0], rec.px1[1], rec.px1[2]};
  uint32_t px2[3] = {rec.px2[0], rec.px2[1], rec.px2[2]};
  uint32_t cl = rec.cidx_lo, ch = rec.cidx_hi;
  uint32_t al = rec.aidx_lo, ah = rec.aidx_hi;
  const int nib = (im == 0) ? a.idx_bits : a.alpha_idx_bits;
  const int nab = (im == 0) ? a.alpha_idx_bits : a.idx_bits;
  // subset loops run to 3 with a guard, so px1/px2 are indexed by constants
#pragma unroll
  for (int s = 0; s < 3; s++) {
    if (s >= ns) continue;
    const uint32_t smask = subset_mask(shape, ns, s);
    const int anchor = anchor_of(s, shape, ns);
    if (idx_at(cl, ch, anchor) >> (nib - 1)) {
      const uint32_t t = px1[s];
      px1[s] = px2[s];
      px2[s] = t;
      invert_idx(cl, ch, smask, (1u << nib) - 1);
      if (a.has_rot) invert_idx(al, ah, 0xFFFFu, (1u << nab) - 1);
    }
    if (a.has_rot && (idx_at(al, ah, anchor) >> (nab - 1)) > 0) {
      const uint32_t a1 = px1[s] >> 24, a2 = px2[s] >> 24;
      px1[s] = (px1[s] & 0x00FFFFFFu) | (a2 << 24);
      px2[s] = (px2[s] & 0x00FFFFFFu) | (a1 << 24);
      invert_idx(al, ah, 0xFFFFu, (1u << nab) - 1);
    }
  }
  const int cp = a.color_prec, ap = a.alpha_prec;
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int s = 0; s < 3; s++) {
      if (s >= ns) continue;
      w.put(chan(px1[s], c) >> (8 - cp), cp);
      w.put(chan(px2[s], c) >> (8 - cp), cp);
    }
  if (ap) {
#pragma unroll
    for (int s = 0; s < 3; s++) {
      if (s >= ns) continue;
      w.put(chan(px1[s], 3) >> (8 - ap), ap);
      w.put(chan(px2[s], 3) >> (8 - ap), ap);
    }
  }
  if (a.pbit_type != PB_NONE) {
#pragma unroll
    for (int s = 0; s < 3; s++) {
      if (s >= ns) continue;
      const int pbc = (rec.meta >> (2 * s)) & 3;
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

// dst == nullptr writes the block into the job's output; otherwise into
// dst[rel] (the quality-0 copy the guard compares against).
__global__ void __launch_bounds__(256) pack_kernel(BlockRange r, Bc7Scratch scr, Work w,
                                                   uint4 *dst) {
  const uint32_t rel = blockIdx.x * blockDim.x + threadIdx.x;
  if (rel >= r.num_blocks) return;
  const uint32_t b = r.first_block + rel;
  const uint32_t sel = w.sel[rel];
  BitWriter bw;
  if (sel_cls(sel) == 0) {
    const uint32_t bx = b % r.blocks_wide, by = b / r.blocks_wide;
    const uint32_t pixel = r.img[(size_t)(by * 4) * r.width + bx * 4];
    const uint32_t wm = (scr.chunk_prefix[rel / 256] + scr.local_prefix[rel]) % 9;
    pack_solid(pixel, wm, bw);
  } else if (sel_cls(sel) == 1) {
    bw.put(1u << 6, 7);  // WriteTransparentBlock (Compressor.cpp:1416-1421)
  } else {
    // the slots hold the candidates in CompressClusters' order: keep the
    // first strict minimum
    const CandRec *rc = &w.rec[(size_t)rel * kMaxCands];
    const int count = sel_count(sel);
    int bc = -1;
    if (w.nu_total) {
      // the nonuniform metric's double totals (the reference's best_err
      // starts at DBL_MAX, Compressor.cpp:1756)
      const double *tn = w.nu_total + (size_t)rel * kMaxCands;
      double best = DBL_MAX;
      for (int i = 0; i < count; i++) {
        if (tn[i] < best) {
          best = tn[i];
          bc = i;
        }
      }
    } else {
      uint32_t best = 0xFFFFFFFFu;
      for (int i = 0; i < count; i++) {
        const uint32_t e = rc[i].err;
        if (e < best) {
          best = e;
          bc = i;
        }
      }
    }
    if (bc >= 0) {
      Cand k;
      cand_from_desc(w.cdesc[(size_t)rel * kMaxCands + bc], k);
      pack_mode(k.mode, k.shape, rc[bc], bw);
    }
  }
  uint4 v;
  v.x = (uint32_t)bw.lo;
  v.y = (uint32_t)(bw.lo >> 32);
  v.z = (uint32_t)bw.hi;
  v.w = (uint32_t)(bw.hi >> 32);
  if (dst) dst[rel] = v;
  else reinterpret_cast<uint4 *>(r.out)[b] = v;
}

// Quality guard (no counterpart in the reference): where the annealed block
// decodes (reference decoder) to a larger premultiplied RGB error than the
// quality-0 block, the quality-0 block is written instead.  The error is
// sum((a*c - a'*c')^2) over pixels and RGB, Image::ComputePSNR's per-pixel
// term (Base/src/Image.cpp:205-255) scaled by 255^2 to stay an exact integer.
__device__ __forceinline__ uint64_t premul_err(const uint32_t src[16], const uint32_t dec[16]) {
  uint64_t e = 0;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int a = (int)(src[i] >> 24), b = (int)(dec[i] >> 24);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const int d = a * (int)chan(src[i], c) - b * (int)chan(dec[i], c);
      e += (uint64_t)((int64_t)d * d);
    }
  }
  return e;
}

__global__ void __launch_bounds__(256) guard_kernel(BlockRange r, const uint4 *q0) {
  const uint32_t rel = blockIdx.x * blockDim.x + threadIdx.x;
  if (rel >= r.num_blocks) return;
  const uint32_t b = r.first_block + rel;
  uint4 *out = reinterpret_cast<uint4 *>(r.out);
  const uint4 a = out[b], z = q0[rel];
  if (a.x == z.x && a.y == z.y && a.z == z.z && a.w == z.w) return;
  uint32_t px[16], da[16], dz[16];
  load_block(r, rel, px);
  bc7dec::decode_block(a, da);
  bc7dec::decode_block(z, dz);
  if (premul_err(px, dz) < premul_err(px, da)) out[b] = z;
}

// The learned search chooses among a block's candidates by decoded error:
// each candidate's block is packed and decoded with the reference decoder,
// and the smallest premul_err wins (ties to the lower slot).  Pass 0 (before
// annealing) keeps the quality-0 winner and its error; the final pass writes
// its winner unless the quality-0 one decodes strictly better, so no block is
// worse than at quality 0.  The errors come from cand_error_kernel, one
// candidate per thread; pick_kernel then chooses per block.

// Candidate decoded errors into cerr[id].  Pass 0 walks the type queues (a
// wave packs one mode) and also fills the slot of a mode-6 second shape
// entry, whose set-up copied the first's record; the final pass walks the
// annealing list and re-scores only the records annealing rewrote.
// Grid-stride: the item counts are device counters.  Four waves per SIMD
// (128 VGPRs, no spills) hide the dependent loads of each item better than
// the three its natural allocation allows.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
cand_error_kernel(BlockRange r, Work w, unsigned long long *cerr, int final_pass) {
  uint32_t total = 0;
  if (final_pass) {
    total = w.counters[18];
  } else {
#pragma unroll
    for (int q = 0; q < kNumTypes; q++) total += w.counters[q];
  }
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
    uint32_t id;
    if (final_pass) {
      id = w.aid[w.alist[t]];
    } else {
      uint32_t base = 0, item = 0;
      int type = kNumTypes - 1;
#pragma unroll
      for (int q = 0; q < kNumTypes; q++) {
        const uint32_t c = w.counters[q];
        if (t >= base && t < base + c) {
          type = q;
          item = t - base;
        }
        base += c;
      }
      id = w.queue[(size_t)type * w.qcap + item];
    }
    const CandRec rc = w.rec[id];
    if (final_pass && !(rc.meta & kMetaAnnealed)) continue;
    const uint32_t rel = id_block(id);
    Cand k;
    cand_from_desc(w.cdesc[id], k);
    BitWriter bw;
    pack_mode(k.mode, k.shape, rc, bw);
    const uint4 blk = make_uint4((uint32_t)bw.lo, (uint32_t)(bw.lo >> 32), (uint32_t)bw.hi,
                                 (uint32_t)(bw.hi >> 32));
    uint32_t px[16], dec[16];
    load_block(r, rel, px);
    bc7dec::decode_block(blk, dec);
    const unsigned long long e = premul_err(px, dec);
    cerr[id] = e;
    if (!final_pass && k.mode == 6) {
      const uint32_t c = id_slot(id);
      if (c + 1 < (uint32_t)sel_count(w.sel[rel])) {
        const uint32_t d1 = w.cdesc[id + 1];
        if ((d1 & 7) == 6 && ((d1 >> 12) & 7) == 1) cerr[id + 1] = e;
      }
    }
  }
}

// Per block: the smallest candidate error (first slot on ties), packed.
// Pass 0 stores the quality-0 winner and its error; the final pass writes
// the block (solid and transparent blocks as pack_kernel does), keeping the
// quality-0 winner when it decodes strictly better (q0err, if given).
//
// A full-search block (kSelFull) chooses as the full search does: the first
// strict minimum of the internal errors (CompressClusters, Compressor.cpp:
// 1752-1817), and with the guard on the final pass keeps the quality-0 block
// when it decodes strictly better (guard_kernel's rule).
__global__ void __launch_bounds__(256) pick_kernel(BlockRange r, Bc7Scratch scr, Work w,
                                                   uint4 *q0blk, unsigned long long *q0err,
                                                   const unsigned long long *cerr,
                                                   int final_pass, int q0_guard) {
  const uint32_t rel = blockIdx.x * blockDim.x + threadIdx.x;
  if (rel >= r.num_blocks) return;
  const uint32_t b = r.first_block + rel;
  const uint32_t sel = w.sel[rel];
  uint4 best = make_uint4(0, 0, 0, 0);
  if (sel_cls(sel) != 2) {
    if (!final_pass) return;
    BitWriter bw;
    if (sel_cls(sel) == 0) {
      const uint32_t bx = b % r.blocks_wide, by = b / r.blocks_wide;
      const uint32_t pixel = r.img[(size_t)(by * 4) * r.width + bx * 4];
      const uint32_t wm = (scr.chunk_prefix[rel / 256] + scr.local_prefix[rel]) % 9;
      pack_solid(pixel, wm, bw);
    } else {
      bw.put(1u << 6, 7);  // WriteTransparentBlock (Compressor.cpp:1416-1421)
    }
    best = make_uint4((uint32_t)bw.lo, (uint32_t)(bw.lo >> 32), (uint32_t)bw.hi,
                      (uint32_t)(bw.hi >> 32));
  } else {
    unsigned long long best_err = ~0ull;
    int best_i = -1;
    const int count = sel_count(sel);
    if (sel_full(sel)) {
      uint32_t best_int = 0xFFFFFFFFu;
      for (int i = 0; i < count; i++) {
        const uint32_t e = w.rec[(size_t)rel * kMaxCands + i].err;
        if (e < best_int) {
          best_int = e;
          best_i = i;
        }
      }
      if (best_i >= 0) best_err = cerr[(size_t)rel * kMaxCands + best_i];
    } else {
      for (int i = 0; i < count; i++) {
        const unsigned long long e = cerr[(size_t)rel * kMaxCands + i];
        if (e < best_err) {
          best_err = e;
          best_i = i;
        }
      }
    }
    if (best_i >= 0) {
      Cand k;
      cand_from_desc(w.cdesc[(size_t)rel * kMaxCands + best_i], k);
      BitWriter bw;
      pack_mode(k.mode, k.shape, w.rec[(size_t)rel * kMaxCands + best_i], bw);
      best = make_uint4((uint32_t)bw.lo, (uint32_t)(bw.lo >> 32), (uint32_t)bw.hi,
                        (uint32_t)(bw.hi >> 32));
    }
    if (!final_pass) {
      q0blk[rel] = best;
      q0err[rel] = best_err;
      return;
    }
    if (q0err && (q0_guard || !sel_full(sel)) && q0err[rel] < best_err) best = q0blk[rel];
  }
  reinterpret_cast<uint4 *>(r.out)[b] = best;
}

// ---------------------------------------------------------------------------
// Exhaustive quality-0 search, the training labels of the candidate ranker:
// every mode with every partition shape, rotation and index mode (the 285
// valid outputs of the ranking network) for every block.  Each candidate is
// fitted, packed and decoded with the reference decoder, and the one with
// the smallest decoded error (premul_err) wins: the encoder's internal error
// is not the decoded error for the rotation modes (stale statistics, the
// reference decoder's mode-4 index-mode-1 reading) and the single-colour
// estimate.  Ties go to the mode tried first by CompressClusters
// (0,2,1,3,7,4,5,6), then to the lower slot.  One launch per mode, so that a
// wave runs one code path; lanes take consecutive slots of one block.
// ---------------------------------------------------------------------------
__constant__ int kModePos[8] = {0, 2, 1, 3, 5, 6, 7, 4};

__device__ __forceinline__ void slot_cand(int mode, int slot, Cand &k) {
  k.mode = mode;
  k.entry = 0;
  k.nsub = kModes[mode].subsets;
  k.shape = k.nsub > 1 ? slot : 0;
  k.rot = mode == 4 ? (slot >> 1) : (mode == 5 ? slot : 0);
  k.im = mode == 4 ? (slot & 1) : 0;
}

__global__ void __launch_bounds__(256) label_fill_kernel(unsigned long long *best, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) best[i] = ~0ull;
}

// ERRORS: every candidate's decoded error to errs (n x 512 floats, logit
// order) instead of the per-block minimum.
template <bool ERRORS>
__global__ void __launch_bounds__(64) label_kernel(BlockRange r, int mode, uint32_t nslots,
                                                   const uint32_t *sctab,
                                                   unsigned long long *best, float *errs) {
  __shared__ uint32_t s_px[64 * kPxStride];
  __shared__ float s_km[16 * 5 * 64];  // k-means sums and counts (up to 16 buckets)
  const uint64_t items = (uint64_t)r.num_blocks * nslots;
  for (uint64_t item = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; item < items;
       item += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t rel = (uint32_t)(item / nslots);
    const int slot = (int)(item % nslots);
    Cand k;
    slot_cand(mode, slot, k);
    uint32_t *px = s_px + threadIdx.x * kPxStride;  // this lane's row only
    {
      uint32_t v[16];
      load_block(r, rel, v);
#pragma unroll
      for (int i = 0; i < 16; i++) px[i] = v[i];
    }
    CandRec rec;
    AState st;
    uint32_t total = 0;
    fit_candidate<-1>(k, px, 0, sctab, rec, st, total, nullptr, nullptr,
                      s_km + threadIdx.x);
    BitWriter bw;
    pack_mode(k.mode, k.shape, rec, bw);
    const uint4 blk = make_uint4((uint32_t)bw.lo, (uint32_t)(bw.lo >> 32), (uint32_t)bw.hi,
                                 (uint32_t)(bw.hi >> 32));
    uint32_t dec[16];
    bc7dec::decode_block(blk, dec);
    const unsigned long long err = premul_err(px, dec);  // < 2^38
    if constexpr (ERRORS) {
      errs[(size_t)rel * 512 + mode * 64 + slot] = (float)err;
    } else {
      const unsigned long long key =
          (err << 16) | (unsigned long long)(kModePos[mode] * 64 + slot);
      atomicMin(&best[rel], key);
    }
  }
}

__global__ void __launch_bounds__(256) label_finish_kernel(const unsigned long long *best,
                                                           uint32_t n, int32_t *labels,
                                                           unsigned long long *errors) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long key = best[i];
  const int rank = (int)(key & 0xFFFFu);
  labels[i] = kOrder[rank >> 6] * 64 + (rank & 63);
  if (errors) errors[i] = key >> 16;
}

// Batches: global blocks [g0, g0 + n) of the BC7 group to a block array
// (16 pixels each), and the encoded blocks back to their surfaces.
__global__ void __launch_bounds__(256) gather_kernel(const SurfDesc *surfs,
                                                     const uint32_t *block_surf, uint32_t g0,
                                                     uint32_t n, uint32_t *blocks) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;  // block * 4 + row
  const uint32_t k = t >> 2, row = t & 3;
  if (k >= n) return;
  const SurfDesc sd = surfs[block_surf[g0 + k]];
  const uint32_t b = g0 + k - sd.first;
  const uint32_t bx = b % sd.blocks_wide, by = b / sd.blocks_wide;
  const uint4 v = *reinterpret_cast<const uint4 *>(sd.img + (size_t)(by * 4 + row) * sd.width +
                                                   bx * 4);
  reinterpret_cast<uint4 *>(blocks)[(size_t)k * 4 + row] = v;
}

__global__ void __launch_bounds__(256) scatter_kernel(const SurfDesc *surfs,
                                                      const uint32_t *block_surf, uint32_t g0,
                                                      uint32_t n, const uint8_t *enc) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const SurfDesc sd = surfs[block_surf[g0 + k]];
  reinterpret_cast<uint4 *>(sd.out)[g0 + k - sd.first] =
      reinterpret_cast<const uint4 *>(enc)[k];
}

__global__ void __launch_bounds__(256) fill_kernel(uint32_t *p, uint32_t n, uint32_t v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// The annealing states appended so far (before the rotation modes' set-up).
__global__ void mark_kernel(uint32_t *counters) { counters[19] = counters[16]; }

__global__ void reset_kernel(uint32_t *counters, int n) {
  const int i = threadIdx.x;
  if (i < n) counters[i] = 0;
}

// Solid-block flags and in-chunk exclusive prefix (chunks of 256 blocks).
__global__ void __launch_bounds__(256) solid_scan_kernel(BlockRange r, Bc7Scratch scr) {
  __shared__ uint32_t s_wave[4];
  const uint32_t t = threadIdx.x;
  const uint32_t rel = blockIdx.x * 256 + t;
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
int g_num_cus[64];
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
  static uint8_t quant[9 * 3 * 256];
  for (int bits = 0; bits <= 8; bits++) {
    const uint32_t mask = bits == 0 ? 0u : ((0xFFu << (8 - bits)) & 0xFFu);
    for (int pb = -1; pb <= 1; pb++)
      for (uint32_t v = 0; v < 256; v++)
        quant[(bits * 3 + pb + 1) * 256 + v] = (uint8_t)bc7::quant_channel(v, mask, pb);
  }
  e = hipMemcpyToSymbol(HIP_SYMBOL(bc7::kQuant), quant, sizeof(quant));
  if (e != hipSuccess) return e;
  e = hipMemcpyToSymbol(HIP_SYMBOL(bc7::kDxt6Tab), d6, sizeof(d6));
  if (e != hipSuccess) return e;
  int cus = 0;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  if (e != hipSuccess) return e;
  g_num_cus[device] = cus > 0 ? cus : 1;
  // the annealing step computes index weights by multiply-shift; it must
  // reproduce the interpolation table entry for entry
  static const uint8_t kWeights[3][16] = {
      {0, 21, 43, 64},
      {0, 9, 18, 27, 37, 46, 55, 64},
      {0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64}};
  for (int t = 0; t < 3; t++) {
    const int nb = 4 << t;
    for (int j = 0; j < nb; j++)
      if (bc7::weight_of(bc7::weight_fn(nb), (uint32_t)j) != kWeights[t][j])
        return hipErrorInvalidValue;
  }
  g_tables_uploaded[device] = 1;
  return hipSuccess;
}

void launch_anneal(uint32_t grid, const BlockRange &r, const Bc7Params &p, const bc7::Work &w,
                   hipStream_t stream) {
  if (p.work_counts && !w.nu_total)  // the counting instance (both searches)
    hipLaunchKernelGGL((bc7::anneal_kernel<4, 3, false, true>), dim3(grid), dim3(64), 0, stream,
                       r, p, w);
  else if (w.nu_total)  // nonuniform metric (the full search or given selections)
    hipLaunchKernelGGL((bc7::anneal_kernel<4, 1, true>), dim3(grid), dim3(64), 0, stream, r, p,
                       w);
  else if (w.alist)  // learned search: four steps between refills
    hipLaunchKernelGGL((bc7::anneal_kernel<4, 3>), dim3(grid), dim3(64), 0, stream, r, p, w);
  else if (p.anneal_unroll >= 4)
    hipLaunchKernelGGL((bc7::anneal_kernel<4, 3>), dim3(grid), dim3(64), 0, stream, r, p, w);
  else if (p.anneal_unroll == 2)
    hipLaunchKernelGGL((bc7::anneal_kernel<2, 3>), dim3(grid), dim3(64), 0, stream, r, p, w);
  else
    hipLaunchKernelGGL((bc7::anneal_kernel<1, 3>), dim3(grid), dim3(64), 0, stream, r, p, w);
}

hipError_t fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(bc7::fill_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, p, n, v);
  return hipGetLastError();
}

hipError_t bc7_prepare() {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = bc7_upload_tables(dev);
  if (e != hipSuccess) return e;
  // the device state (single-colour table, watermark counter) is created on
  // first use
  if (!device_wm_counter(dev)) return hipErrorNotInitialized;
  return bc7_single_color_table(dev) ? hipSuccess : hipErrorNotInitialized;
}

hipError_t bc7_gather(const SurfDesc *surfs, const uint32_t *block_surf, uint32_t g0, uint32_t n,
                      uint32_t *blocks, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(bc7::gather_kernel, dim3((n * 4 + 255) / 256), dim3(256), 0, stream, surfs,
                     block_surf, g0, n, blocks);
  return hipGetLastError();
}

hipError_t bc7_scatter(const SurfDesc *surfs, const uint32_t *block_surf, uint32_t g0, uint32_t n,
                       const uint8_t *enc, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(bc7::scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, surfs,
                     block_surf, g0, n, enc);
  return hipGetLastError();
}

size_t bc7_work_bytes(uint32_t num_blocks, bool learned, bool nonuniform) {
  const size_t n = num_blocks;
  // counters, sel, cdesc, rec, queues, astate, q0, alignment slack
  size_t bytes = 64 * 4 + n * 4 + n * bc7::kMaxCands * 2 +
                 n * bc7::kMaxCands * sizeof(bc7::CandRec) +
                 n * bc7::kNumTypes * bc7::kMaxPerType * 4 +
                 n * bc7::kMaxCands * sizeof(bc7::AState) + n * 16 + 8 * 256;
  if (learned) {
    // q0 errors, candidate errors, annealing list, fallback list
    bytes += n * 8 + n * bc7::kMaxCands * (8 + 4 + 4) + n * 4 +
             (n * bc7::kMaxCands / bc7::kListChunk + 1) * 4 + 11 * 256;
  }
  if (nonuniform) bytes += n * bc7::kMaxCands * 8 + 256;  // candidate totals
  return bytes;
}

hipError_t bc7_encode(const BlockRange &r, const Bc7Params &p, const Bc7Scratch &s, void *work,
                      hipStream_t stream) {
  if (r.num_blocks == 0) return hipSuccess;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = bc7_upload_tables(dev);
  if (e != hipSuccess) return e;
  const uint32_t *sct = bc7_single_color_table(dev);
  if (!sct) return hipErrorOutOfMemory;
  const uint32_t n = r.num_blocks;
  // carve the work area
  uint8_t *base = reinterpret_cast<uint8_t *>(work);
  auto carve = [&](size_t bytes) {
    uint8_t *q = base;
    base += (bytes + 255) & ~(size_t)255;
    return q;
  };
  bc7::Work w;
  const size_t nc = (size_t)n * bc7::kMaxCands;
  w.counters = reinterpret_cast<uint32_t *>(carve(64 * 4));
  w.sel = reinterpret_cast<uint32_t *>(carve((size_t)n * 4));
  w.cdesc = reinterpret_cast<uint16_t *>(carve(nc * 2));
  w.rec = reinterpret_cast<bc7::CandRec *>(carve(nc * sizeof(bc7::CandRec)));
  w.qcap = (uint32_t)((size_t)n * bc7::kMaxPerType);
  w.queue = reinterpret_cast<uint32_t *>(carve((size_t)w.qcap * bc7::kNumTypes * 4));
  w.astate = reinterpret_cast<bc7::AState *>(carve(nc * sizeof(bc7::AState)));
  w.astate_cap = (uint32_t)nc;
  uint4 *q0 = reinterpret_cast<uint4 *>(carve((size_t)n * 16));
  const bool learned = p.ranker != nullptr;
  unsigned long long *q0err = nullptr;
  uint32_t *fallback = nullptr;
  unsigned long long *cerr = nullptr;
  w.alist = nullptr;
  w.aid = nullptr;
  w.nu_total = nullptr;
  // the nonuniform metric runs the reference search only (the host entry
  // points reject it with the ranker or statistics)
  const bool nu = p.nonuniform != 0 && !learned && !p.stats;
  if (nu) w.nu_total = reinterpret_cast<double *>(carve(nc * 8));
  uint32_t *alist = nullptr, *chunk_count = nullptr;
  if (learned) {
    q0err = reinterpret_cast<unsigned long long *>(carve((size_t)n * 8));
    cerr = reinterpret_cast<unsigned long long *>(carve(nc * 8));
    alist = reinterpret_cast<uint32_t *>(carve(nc * 4));
    w.aid = reinterpret_cast<uint32_t *>(carve(nc * 4));
    chunk_count = reinterpret_cast<uint32_t *>(carve((nc / bc7::kListChunk + 1) * 4));
    fallback = reinterpret_cast<uint32_t *>(carve((size_t)n * 4));
  }
  uint32_t *fallback_count = w.counters + 20;
  const uint32_t nchunks = (n + 255) / 256;
  const uint32_t ngroups = (n + 15) / 16;
  // persistent workgroups of one wave, no more than the candidates could use
  uint32_t persistent = (uint32_t)g_num_cus[dev] * (p.wg_per_cu ? p.wg_per_cu : 16u);
  const uint64_t max_waves = ((uint64_t)n * bc7::kMaxCands + 63) / 64;
  if (persistent > max_waves) persistent = (uint32_t)max_waves;
  const dim3 pack_grid((n + 255) / 256);
  // candidate-error kernel: grid-stride, a few workgroups per CU
  const uint32_t err_grid = (uint32_t)g_num_cus[dev] * 8u;
  const bool guard = p.quality > 0 && p.q0_guard;
  int t = timing_begin(stream);
  hipLaunchKernelGGL(bc7::reset_kernel, dim3(1), dim3(64), 0, stream, w.counters, 64);
  hipLaunchKernelGGL(bc7::solid_scan_kernel, dim3(nchunks), dim3(256), 0, stream, r, s);
  hipLaunchKernelGGL(bc7::chunk_scan_kernel, dim3(1), dim3(1024), 0, stream, s);
  timing_end(t, kTagBc7Scan, stream);
  if (learned) {
    // rank every block and keep the confident top-k, then run BoxSelection
    // over the blocks the ranker was not sure about
    const Ranker &rk = *p.ranker;
    t = timing_begin(stream);
    hipLaunchKernelGGL(bc7::rank_select_kernel, dim3((n + mlp::kRankRows - 1) / mlp::kRankRows),
                       dim3(256), 0, stream, r, p, w, (const uint16_t *)rk.w1_oi,
                       (const uint16_t *)rk.w2_oi, (const uint16_t *)rk.w3_oi,
                       (const float *)rk.master, fallback, fallback_count);
    timing_end(t, kTagBc7Rank, stream);
    t = timing_begin(stream);
    hipLaunchKernelGGL(bc7::select_kernel<bc7::kSelBox>, dim3(ngroups), dim3(256), 0, stream,
                       r, p, w, (const uint32_t *)fallback, (const uint32_t *)fallback_count);
    timing_end(t, kTagBc7Select, stream);
  } else if (p.stats) {
    t = timing_begin(stream);
    hipLaunchKernelGGL(bc7::select_kernel<bc7::kSelStats>, dim3(ngroups), dim3(256), 0, stream,
                       r, p, w, (const uint32_t *)nullptr, (const uint32_t *)nullptr);
    timing_end(t, kTagBc7Select, stream);
  } else if (p.selections) {
    t = timing_begin(stream);
    hipLaunchKernelGGL(bc7::select_kernel<bc7::kSelGiven>, dim3(ngroups), dim3(256), 0, stream,
                       r, p, w, (const uint32_t *)nullptr, (const uint32_t *)nullptr);
    timing_end(t, kTagBc7Select, stream);
  } else if (nu) {
    t = timing_begin(stream);
    hipLaunchKernelGGL((bc7::select_kernel<bc7::kSelBox, true>), dim3(ngroups), dim3(256), 0,
                       stream, r, p, w, (const uint32_t *)nullptr, (const uint32_t *)nullptr);
    timing_end(t, kTagBc7Select, stream);
  } else {
    t = timing_begin(stream);
    hipLaunchKernelGGL(bc7::select_kernel<bc7::kSelBox>, dim3(ngroups), dim3(256), 0, stream,
                       r, p, w, (const uint32_t *)nullptr, (const uint32_t *)nullptr);
    timing_end(t, kTagBc7Select, stream);
  }
  t = timing_begin(stream);
  // annealing-state order: modes 0, 2, 3, 7, 6, 1, then the rotation modes 4
  // and 5 (the mark between them records where they begin)
  if (nu) {
    hipLaunchKernelGGL((bc7::setup_kernel<0, true>), dim3(persistent), dim3(64), 0, stream, r,
                       p, w, sct);
    hipLaunchKernelGGL((bc7::setup_kernel<2, true>), dim3(persistent), dim3(64), 0, stream, r,
                       p, w, sct);
    hipLaunchKernelGGL((bc7::setup_kernel<3, true>), dim3(persistent), dim3(64), 0, stream, r,
                       p, w, sct);
    hipLaunchKernelGGL((bc7::setup_kernel<7, true>), dim3(persistent), dim3(64), 0, stream, r,
                       p, w, sct);
    hipLaunchKernelGGL((bc7::setup_kernel<6, true>), dim3(persistent), dim3(64), 0, stream, r,
                       p, w, sct);
    hipLaunchKernelGGL((bc7::setup_kernel<1, true>), dim3(persistent), dim3(64), 0, stream, r,
                       p, w, sct);
    hipLaunchKernelGGL(bc7::mark_kernel, dim3(1), dim3(1), 0, stream, w.counters);
    hipLaunchKernelGGL((bc7::setup_kernel<4, true>), dim3(persistent), dim3(64), 0, stream, r,
                       p, w, sct);
    hipLaunchKernelGGL((bc7::setup_kernel<5, true>), dim3(persistent), dim3(64), 0, stream, r,
                       p, w, sct);
  } else {
    hipLaunchKernelGGL(bc7::setup_kernel<0>, dim3(persistent), dim3(64), 0, stream, r, p, w, sct);
    hipLaunchKernelGGL(bc7::setup_kernel<2>, dim3(persistent), dim3(64), 0, stream, r, p, w, sct);
    hipLaunchKernelGGL(bc7::setup_kernel<3>, dim3(persistent), dim3(64), 0, stream, r, p, w, sct);
    hipLaunchKernelGGL(bc7::setup_kernel<7>, dim3(persistent), dim3(64), 0, stream, r, p, w, sct);
    hipLaunchKernelGGL(bc7::setup_kernel<6>, dim3(persistent), dim3(64), 0, stream, r, p, w, sct);
    hipLaunchKernelGGL(bc7::setup_kernel<1>, dim3(persistent), dim3(64), 0, stream, r, p, w, sct);
    hipLaunchKernelGGL(bc7::mark_kernel, dim3(1), dim3(1), 0, stream, w.counters);
    hipLaunchKernelGGL(bc7::setup_kernel<4>, dim3(persistent), dim3(64), 0, stream, r, p, w, sct);
    hipLaunchKernelGGL(bc7::setup_kernel<5>, dim3(persistent), dim3(64), 0, stream, r, p, w, sct);
  }
  timing_end(t, kTagBc7Setup, stream);
  if (learned) {
    if (p.quality > 0) {
      t = timing_begin(stream);
      hipLaunchKernelGGL(bc7::cand_error_kernel, dim3(err_grid), dim3(256), 0, stream, r, w,
                         cerr, 0);
      hipLaunchKernelGGL(bc7::pick_kernel, pack_grid, dim3(256), 0, stream, r, s, w, q0, q0err,
                         (const unsigned long long *)cerr, 0, (int)p.q0_guard);
      timing_end(t, kTagBc7Pack, stream);
      bc7::Work wa = w;
      wa.alist = alist;
      t = timing_begin(stream);
      // ordered compaction of the annealing list (grid covers the largest list)
      const uint32_t nlist = (uint32_t)((nc + bc7::kListChunk - 1) / bc7::kListChunk);
      hipLaunchKernelGGL(bc7::list_count_kernel, dim3(nlist), dim3(256), 0, stream,
                         (const uint32_t *)w.aid, (const uint32_t *)w.counters,
                         (const uint32_t *)w.sel, cerr, q0err, chunk_count);
      hipLaunchKernelGGL(bc7::list_scan_kernel, dim3(1), dim3(256), 0, stream, w.counters,
                         chunk_count);
      hipLaunchKernelGGL(bc7::list_write_kernel, dim3(nlist), dim3(256), 0, stream,
                         (const uint32_t *)w.aid, (const uint32_t *)w.counters,
                         (const uint32_t *)w.sel, cerr, q0err, (const uint32_t *)chunk_count,
                         alist);
      launch_anneal(persistent, r, p, wa, stream);
      timing_end(t, kTagBc7Anneal, stream);
    }
    t = timing_begin(stream);
    bc7::Work wf = w;
    wf.alist = alist;
    // quality 0: every candidate's error; otherwise the annealed ones
    hipLaunchKernelGGL(bc7::cand_error_kernel, dim3(err_grid), dim3(256), 0, stream, r, wf, cerr,
                       p.quality > 0 ? 1 : 0);
    hipLaunchKernelGGL(bc7::pick_kernel, pack_grid, dim3(256), 0, stream, r, s, w, q0,
                       p.quality > 0 ? q0err : (unsigned long long *)nullptr,
                       (const unsigned long long *)cerr, 1, (int)p.q0_guard);
    timing_end(t, kTagBc7Pack, stream);
    return hipGetLastError();
  }
  if (guard) {
    t = timing_begin(stream);
    hipLaunchKernelGGL(bc7::pack_kernel, pack_grid, dim3(256), 0, stream, r, s, w, q0);
    timing_end(t, kTagBc7Pack, stream);
  }
  if (p.quality > 0) {
    t = timing_begin(stream);
    launch_anneal(persistent, r, p, w, stream);
    timing_end(t, kTagBc7Anneal, stream);
  }
  t = timing_begin(stream);
  hipLaunchKernelGGL(bc7::pack_kernel, pack_grid, dim3(256), 0, stream, r, s, w,
                     (uint4 *)nullptr);
  timing_end(t, kTagBc7Pack, stream);
  if (guard) {
    t = timing_begin(stream);
    hipLaunchKernelGGL(bc7::guard_kernel, pack_grid, dim3(256), 0, stream, r, q0);
    timing_end(t, kTagBc7Guard, stream);
  }
  if (p.stats)
    hipLaunchKernelGGL(bc7::stats_mode_kernel, pack_grid, dim3(256), 0, stream, r, p.stats);
  return hipGetLastError();
}

hipError_t bc7_exhaustive_labels(const BlockRange &r, unsigned long long *best,
                                 int32_t *labels, unsigned long long *errors,
                                 hipStream_t stream) {
  const uint32_t n = r.num_blocks;
  if (n == 0) return hipSuccess;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = bc7_upload_tables(dev);
  if (e != hipSuccess) return e;
  const uint32_t *sct = bc7_single_color_table(dev);
  if (!sct) return hipErrorOutOfMemory;
  hipLaunchKernelGGL(bc7::label_fill_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, best,
                     n);
  static const uint32_t kSlots[8] = {16, 64, 64, 64, 8, 4, 1, 64};
  const int t = timing_begin(stream);
  for (int mode = 0; mode < 8; mode++) {
    const uint64_t items = (uint64_t)n * kSlots[mode];
    uint64_t groups = (items + 63) / 64;
    const uint64_t cap = (uint64_t)g_num_cus[dev] * 64;
    if (groups > cap) groups = cap;
    hipLaunchKernelGGL(bc7::label_kernel<false>, dim3((uint32_t)groups), dim3(64), 0, stream, r,
                       mode, kSlots[mode], sct, best, (float *)nullptr);
  }
  timing_end(t, kTagBc7Labels, stream);
  hipLaunchKernelGGL(bc7::label_finish_kernel, dim3((n + 255) / 256), dim3(256), 0, stream,
                     (const unsigned long long *)best, n, labels, errors);
  return hipGetLastError();
}

hipError_t bc7_candidate_errors(const BlockRange &r, float *errs, hipStream_t stream) {
  const uint32_t n = r.num_blocks;
  if (n == 0) return hipSuccess;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = bc7_upload_tables(dev);
  if (e != hipSuccess) return e;
  const uint32_t *sct = bc7_single_color_table(dev);
  if (!sct) return hipErrorOutOfMemory;
  // logits that are no candidate stay +inf
  const uint64_t words = (uint64_t)n * 512;
  hipLaunchKernelGGL(bc7::fill_kernel, dim3((uint32_t)((words + 255) / 256)), dim3(256), 0, stream,
                     reinterpret_cast<uint32_t *>(errs), (uint32_t)words, 0x7F800000u);
  static const uint32_t kSlots[8] = {16, 64, 64, 64, 8, 4, 1, 64};
  const int t = timing_begin(stream);
  for (int mode = 0; mode < 8; mode++) {
    const uint64_t items = (uint64_t)n * kSlots[mode];
    uint64_t groups = (items + 63) / 64;
    const uint64_t cap = (uint64_t)g_num_cus[dev] * 64;
    if (groups > cap) groups = cap;
    hipLaunchKernelGGL(bc7::label_kernel<true>, dim3((uint32_t)groups), dim3(64), 0, stream, r,
                       mode, kSlots[mode], sct, (unsigned long long *)nullptr, errs);
  }
  timing_end(t, kTagBc7Labels, stream);
  return hipGetLastError();
}

}  // namespace fh
