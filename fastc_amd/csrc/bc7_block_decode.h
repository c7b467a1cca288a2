// BC7 (BPTC) block decoder shared by the decode kernel and the encoder's
// quality guard: BPTCC::Decompress (BPTCEncoder/src/Decompressor.cpp:32-367).
//
// Like the reference, mode 4 blocks with index mode 1 interpolate with the
// per-mode index widths rather than the swapped ones (Decompressor.cpp:263-295),
// so decoded pixels match the reference decoder byte for byte.  With SPEC the
// widths swap with the indices, as the BC7 specification (and hardware)
// decodes such blocks.  Blocks with the reserved mode (no mode bit in the
// first byte) decode to zero.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fh {
namespace bc7dec {

struct Attr {
  int8_t subsets, part_bits, rot_bits, idxm_bits, cprec, aprec, pbit, idx_bits, aidx_bits;
};
// pbit: 0 none, 1 shared per subset, 2 one per endpoint
static __constant__ Attr kAttr[8] = {{3, 4, 0, 0, 4, 0, 2, 3, 0}, {2, 6, 0, 0, 6, 0, 1, 3, 0},
                              {3, 6, 0, 0, 5, 0, 0, 2, 0}, {2, 6, 0, 0, 7, 0, 2, 2, 0},
                              {1, 0, 2, 1, 5, 6, 0, 2, 3}, {1, 0, 2, 0, 7, 8, 0, 2, 2},
                              {1, 0, 0, 0, 7, 7, 2, 4, 0}, {2, 6, 0, 0, 5, 5, 2, 2, 0}};
static __constant__ uint16_t kS2[64] = {
    0xcccc, 0x8888, 0xeeee, 0xecc8, 0xc880, 0xfeec, 0xfec8, 0xec80, 0xc800, 0xffec, 0xfe80,
    0xe800, 0xffe8, 0xff00, 0xfff0, 0xf000, 0xf710, 0x008e, 0x7100, 0x08ce, 0x008c, 0x7310,
    0x3100, 0x8cce, 0x088c, 0x3110, 0x6666, 0x366c, 0x17e8, 0x0ff0, 0x718e, 0x399c, 0xaaaa,
    0xf0f0, 0x5a5a, 0x33cc, 0x3c3c, 0x55aa, 0x9696, 0xa55a, 0x73ce, 0x13c8, 0x324c, 0x3bdc,
    0x6996, 0xc33c, 0x9966, 0x0660, 0x0272, 0x04e4, 0x4e40, 0x2720, 0xc936, 0x936c, 0x39c6,
    0x639c, 0x9336, 0x9cc6, 0x817e, 0xe718, 0xccf0, 0x0fcc, 0x7744, 0xee22};
static __constant__ uint16_t kS3[64][2] = {
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
static __constant__ uint8_t kA2[64] = {
    15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 2,  8,  2,  2, 8,
    8,  15, 2,  8,  2,  2,  8,  8,  2,  2,  15, 15, 6,  8,  2,  8,  15, 15, 2,  8,  2, 2,
    2,  15, 15, 6,  6,  2,  6,  8,  15, 15, 2,  2,  15, 15, 15, 15, 15, 2,  2,  15};
static __constant__ uint8_t kA3[2][64] = {
    {3,  3, 15, 15, 8, 3,  15, 15, 8,  8,  6,  6,  6,  5,  3,  3,  3,  3,  8, 15, 3,  3,
     6,  10, 5, 8,  8, 6,  8,  5,  15, 15, 8,  15, 3,  5,  6,  10, 8,  15, 15, 3, 15, 5,
     15, 15, 15, 15, 3, 15, 5,  5,  5,  8,  5,  10, 5,  10, 8,  13, 15, 12, 3,  3},
    {15, 8,  8,  3,  15, 15, 3,  8,  15, 15, 15, 15, 15, 15, 15, 8,  15, 8,  15, 3,  15, 8,
     15, 8,  3,  15, 6,  10, 15, 15, 10, 8,  15, 3,  15, 10, 10, 8,  9,  10, 6,  15, 8,  15,
     3,  6,  6,  8,  15, 3,  15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 3,  15, 15, 8}};
// Bits [pos, pos + n) of the 128-bit block (n <= 8; pos + n <= 128).
__device__ __forceinline__ uint32_t bits_at(uint64_t lo, uint64_t hi, uint32_t pos, uint32_t n) {
  // shift counts masked to 0..63: the branch not taken never shifts out of range
  const uint64_t v = pos >= 64 ? (hi >> ((pos - 64) & 63u))
                               : ((lo >> (pos & 63u)) | ((hi << 1) << ((63 - pos) & 63u)));
  return (uint32_t)v & ((1u << n) - 1u);
}

// Interpolation weight of index v for n-bit indices, the reference's
// kInterpolationValues (Compressor.cpp:159-168): round(64 v / (2^n - 1)),
// computed as (128 v + m) / (2 m) with a 16-bit reciprocal that is exact for
// these operands; 0 for an index past the table row, which the mode-4 index-mode-1
// quirk reads.
__device__ __forceinline__ uint32_t weight1(uint32_t v, uint32_t n) {
  const uint32_t m = (1u << n) - 1u;
  const uint32_t recip = n == 2 ? 10923u : (n == 3 ? 4682u : 2185u);
  return v <= m ? ((128u * v + m) * recip) >> 16 : 0u;
}

// One channel of the interpolated colour: (e0 * w0 + e1 * w1 + 32) >> 6 with
// w0 = 64 - w1 for a valid index, and both weights zero past the table.
__device__ __forceinline__ uint32_t lerp_chan(uint32_t e0, uint32_t e1, uint32_t w1, bool valid) {
  return valid ? ((e0 * (64u - w1) + e1 * w1 + 32u) >> 6) & 0xFFu : 0u;
}

// Decodes one 16-byte block (little-endian 128-bit stream) into 16 RGBA8
// pixels in raster order.  Every field position is computed from the mode's
// attributes, so lanes holding blocks of different modes run one path and
// nothing is indexed dynamically in private memory.
template <bool SPEC = false>
__device__ inline void decode_block(const uint4 raw, uint32_t px[16]) {
  const uint64_t lo = (uint64_t)raw.x | ((uint64_t)raw.y << 32);
  const uint64_t hi = (uint64_t)raw.z | ((uint64_t)raw.w << 32);
  const uint32_t first = raw.x & 0xFFu;
  if (first == 0) {  // reserved mode
#pragma unroll
    for (int i = 0; i < 16; i++) px[i] = 0;
    return;
  }
  const int mode = __builtin_ctz(first);
  const Attr a = kAttr[mode];
  const uint32_t ns = a.subsets, cp = a.cprec, ap = a.aprec;
  uint32_t pos = mode + 1;
  const uint32_t shape = bits_at(lo, hi, pos, a.part_bits);
  pos += a.part_bits;
  const uint32_t rot = bits_at(lo, hi, pos, a.rot_bits);
  pos += a.rot_bits;
  const uint32_t im = bits_at(lo, hi, pos, a.idxm_bits);
  pos += a.idxm_bits;
  // endpoints: R, G, B of every subset and endpoint, then A, then p-bits
  const uint32_t apos = pos + 6 * ns * cp;
  const uint32_t pbpos = apos + 2 * ns * ap;
  const uint32_t npb = a.pbit == 1 ? ns : (a.pbit == 2 ? 2 * ns : 0);
  const uint32_t cpp = cp + (a.pbit ? 1 : 0), app = ap + (a.pbit ? 1 : 0);
  uint32_t ep[3][2];  // packed RGBA8 endpoints
#pragma unroll
  for (uint32_t s = 0; s < 3; s++)
#pragma unroll
    for (uint32_t e = 0; e < 2; e++) {
      uint32_t w = 0;
      if (s < ns) {
        const uint32_t pb = a.pbit == 1 ? bits_at(lo, hi, pbpos + s, 1)
                                        : (a.pbit == 2 ? bits_at(lo, hi, pbpos + 2 * s + e, 1) : 0u);
#pragma unroll
        for (uint32_t ch = 0; ch < 3; ch++) {
          uint32_t v = (bits_at(lo, hi, pos + ((ch * ns + s) * 2 + e) * cp, cp) << (8 - cp)) & 0xFFu;
          if (a.pbit) v |= (pb << (8 - cpp)) & 0xFFu;
          v |= v >> cpp;
          w |= v << (8 * ch);
        }
        uint32_t al = 0xFFu;
        if (ap) {
          al = (bits_at(lo, hi, apos + (s * 2 + e) * ap, ap) << (8 - ap)) & 0xFFu;
          if (a.pbit) al |= (pb << (8 - app)) & 0xFFu;
          al |= al >> app;
        }
        w |= al << 24;
      }
      ep[s][e] = w;
    }
  // index streams: colour (anchors carry one bit less), then alpha (modes 4, 5)
  const uint32_t ib = a.idx_bits, aib = a.aidx_bits;
  const uint32_t cstart = pbpos + npb;
  const uint32_t astart = cstart + 16 * ib - ns;
  uint32_t m1 = 0, m2 = 0, a1 = 16, a2 = 16;
  if (ns == 2) {
    m1 = kS2[shape];
    a1 = kA2[shape];
  } else if (ns == 3) {
    m1 = kS3[shape][0];
    m2 = kS3[shape][1];
    a1 = kA3[0][shape];
    a2 = kA3[1][shape];
  }
#pragma unroll
  for (uint32_t i = 0; i < 16; i++) {
    const uint32_t skip = (i > 0 ? 1u : 0u) + (i > a1 ? 1u : 0u) + (i > a2 ? 1u : 0u);
    const bool anchor = i == 0 || i == a1 || i == a2;
    uint32_t ci = bits_at(lo, hi, cstart + i * ib - skip, ib - (anchor ? 1u : 0u));
    uint32_t ai = ci;
    if (aib) {
      ai = bits_at(lo, hi, astart + i * aib - (i > 0 ? 1u : 0u), aib - (i == 0 ? 1u : 0u));
      if (im) {
        const uint32_t t = ai;
        ai = ci;
        ci = t;
      }
    }
    // subset: partition masks (3 subsets: second mask picks subset 2)
    const uint32_t sub = ((m1 >> i) & 1u) ? 1u + ((m2 >> i) & 1u) : 0u;
    const uint32_t e0 = sub == 0 ? ep[0][0] : (sub == 1 ? ep[1][0] : ep[2][0]);
    const uint32_t e1 = sub == 0 ? ep[0][1] : (sub == 1 ? ep[1][1] : ep[2][1]);
    // index widths: the reference keeps the per-mode widths after the swap
    const uint32_t cb = SPEC && im ? aib : ib;
    const uint32_t na = SPEC && im ? ib : (aib ? aib : ib);
    const uint32_t wc = weight1(ci, cb);
    const bool cvalid = ci < (1u << cb);
    const uint32_t wa = weight1(ai, na);
    const bool avalid = ai < (1u << na);
    uint32_t pixel = 0;
#pragma unroll
    for (int ch = 0; ch < 3; ch++)
      pixel |= lerp_chan((e0 >> (8 * ch)) & 0xFFu, (e1 >> (8 * ch)) & 0xFFu, wc, cvalid)
               << (8 * ch);
    pixel |= lerp_chan(e0 >> 24, e1 >> 24, wa, avalid) << 24;
    if (rot >= 1) {
      const uint32_t sh = 8 * (rot - 1);
      const uint32_t c = (pixel >> sh) & 0xFFu, al = pixel >> 24;
      pixel = (pixel & ~(0xFFu << sh) & 0x00FFFFFFu) | (al << sh) | (c << 24);
    }
    px[i] = pixel;
  }
}

}  // namespace bc7dec
}  // namespace fh
