// BC7 (BPTC) block decoder shared by the decode kernel and the encoder's
// quality guard: BPTCC::Decompress (BPTCEncoder/src/Decompressor.cpp:32-367).
//
// Like the reference, mode 4 blocks with index mode 1 interpolate with the
// per-mode index widths rather than the swapped ones (Decompressor.cpp:263-295),
// so decoded pixels match the reference decoder byte for byte.  Blocks with
// the reserved mode (no mode bit in the first byte) decode to zero.
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
// weights for 1..4-bit indices; 1-bit entries beyond index 2 are zero, and
// 2-bit entries past 3 are zero, which the mode-4/index-mode-1 path reads.
static __constant__ uint8_t kW0[4][16] = {
    {64, 33, 0},
    {64, 43, 21, 0},
    {64, 55, 46, 37, 27, 18, 9, 0},
    {64, 60, 55, 51, 47, 43, 38, 34, 30, 26, 21, 17, 13, 9, 4, 0}};
static __constant__ uint8_t kW1[4][16] = {
    {0, 31, 64},
    {0, 21, 43, 64},
    {0, 9, 18, 27, 37, 46, 55, 64},
    {0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64}};

struct Reader {
  uint64_t lo, hi;
  int pos;
  __device__ __forceinline__ uint32_t get(int n) {
    if (n <= 0) return 0;
    uint64_t v;
    if (pos >= 64) {
      v = hi >> (pos - 64);
    } else {
      v = lo >> pos;
      if (pos + n > 64) v |= hi << (64 - pos);
    }
    pos += n;
    return (uint32_t)(v & ((1ull << n) - 1ull));
  }
};

__device__ __forceinline__ int subset_of(int i, int shape, int ns) {
  if (ns == 2) return (kS2[shape] >> i) & 1;
  if (ns == 3) {
    if ((kS3[shape][0] >> i) & 1) return 1 + ((kS3[shape][1] >> i) & 1);
    return 0;
  }
  return 0;
}
__device__ __forceinline__ int anchor_of(int s, int shape, int ns) {
  if (s == 1) return ns == 2 ? kA2[shape] : kA3[0][shape];
  if (s == 2) return kA3[1][shape];
  return 0;
}

// Decodes one 16-byte block (little-endian 128-bit stream) into 16 RGBA8
// pixels in raster order.
__device__ inline void decode_block(const uint4 raw, uint32_t px[16]) {
  Reader r;
  r.lo = (uint64_t)raw.x | ((uint64_t)raw.y << 32);
  r.hi = (uint64_t)raw.z | ((uint64_t)raw.w << 32);
  r.pos = 0;
  int mode = 0;
  while (mode < 8 && !r.get(1)) mode++;
  if (mode >= 8) {
#pragma unroll
    for (int i = 0; i < 16; i++) px[i] = 0;
  } else {
    const Attr a = kAttr[mode];
    const int ns = a.subsets;
    int shape = 0, rot = 0, im = 0;
    if (ns > 1) shape = (int)r.get(a.part_bits);
    else if (a.rot_bits) {
      rot = (int)r.get(2);
      if (a.idxm_bits) im = (int)r.get(1);
    }
    uint32_t ep[3][2][4];
    int cp = a.cprec, ap = a.aprec;
    for (int ch = 0; ch < 3; ch++)
      for (int s = 0; s < ns; s++)
        for (int e = 0; e < 2; e++) ep[s][e][ch] = (r.get(cp) << (8 - cp)) & 0xFFu;
    for (int s = 0; s < ns; s++)
      for (int e = 0; e < 2; e++) ep[s][e][3] = ap == 0 ? 0xFFu : (r.get(ap) << (8 - ap)) & 0xFFu;
    if (a.pbit) {
      cp++;
      ap++;
      for (int s = 0; s < ns; s++) {
        uint32_t pb = a.pbit == 1 ? r.get(1) : 0;
        for (int e = 0; e < 2; e++) {
          if (a.pbit == 2) pb = r.get(1);
          for (int ch = 0; ch < 4; ch++)
            ep[s][e][ch] = (ep[s][e][ch] | (pb << (8 - (ch == 3 ? ap : cp)))) & 0xFFu;
        }
      }
    }
    for (int s = 0; s < ns; s++)
      for (int e = 0; e < 2; e++)
        for (int ch = 0; ch < 4; ch++) ep[s][e][ch] |= ep[s][e][ch] >> (ch == 3 ? ap : cp);
    uint32_t cidx[16], aidx[16];
    for (int i = 0; i < 16; i++) {
      const int s = subset_of(i, shape, ns);
      cidx[i] = r.get(anchor_of(s, shape, ns) == i ? a.idx_bits - 1 : a.idx_bits);
    }
    if (a.aidx_bits == 0) {
      for (int i = 0; i < 16; i++) aidx[i] = cidx[i];
    } else {
      for (int i = 0; i < 16; i++) aidx[i] = r.get(i == 0 ? a.aidx_bits - 1 : a.aidx_bits);
      if (im)
        for (int i = 0; i < 16; i++) {
          const uint32_t t = aidx[i];
          aidx[i] = cidx[i];
          cidx[i] = t;
        }
    }
    const int nba = a.aidx_bits, nbc = a.idx_bits;
    for (int i = 0; i < 16; i++) {
      const int s = subset_of(i, shape, ns);
      uint32_t pixel = 0;
      for (int ch = 0; ch < 4; ch++) {
        uint32_t w0, w1;
        if (ch == 3 && nba > 0) {
          w0 = kW0[nba - 1][aidx[i] & 15];
          w1 = kW1[nba - 1][aidx[i] & 15];
        } else {
          w0 = kW0[nbc - 1][cidx[i] & 15];
          w1 = kW1[nbc - 1][cidx[i] & 15];
        }
        pixel |= (((ep[s][0][ch] * w0 + ep[s][1][ch] * w1) + 32) >> 6 & 0xFFu) << (8 * ch);
      }
      if (rot >= 1) {
        const uint32_t sh = 8 * (rot - 1);
        const uint32_t c = (pixel >> sh) & 0xFFu, al = pixel >> 24;
        pixel = (pixel & ~(0xFFu << sh) & 0x00FFFFFFu) | (al << sh) | (c << 24);
      }
      px[i] = pixel;
    }
  }
}

}  // namespace bc7dec
}  // namespace fh
