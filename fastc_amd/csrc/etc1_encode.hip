// ETC1 block encoder for gfx950.
//
// FasTC encodes ETC1 with rg_etc1 at cLowQuality without dithering
// (ETCEncoder/src/Compressor.cpp:26-54).  This kernel reproduces that encoder
// bit for bit (ETCEncoder/src/rg_etc1.cpp):
//   * solid blocks (all 16 RGBA words equal, alpha included) through the
//     exact-match configuration lists and the inverse lookup table
//     (pack_etc1_block_solid_color, :1951-2033);
//   * otherwise 2 flips x {555+delta, 444} x 2 subblocks, each subblock fit by
//     etc1_optimizer with scan delta {0} and at most two refinement trials
//     (compute, :1483-1625; init, :1627-1672; evaluate_solution_fast,
//     :1767-1885), stopping a candidate once its error reaches the best one
//     (:2352), then the selector bit packing (:2364-2449).
// One lane per block; pixels are loaded as four 16-byte rows.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.h"

namespace fh {
namespace etc {

__constant__ int kInten[8][4] = {{-8, -2, 2, 8},       {-17, -5, 5, 17},   {-29, -9, 9, 29},
                                 {-42, -13, 13, 42},   {-60, -18, 18, 60}, {-80, -24, 24, 80},
                                 {-106, -33, 33, 106}, {-183, -47, 47, 183}};
__constant__ uint8_t kSelToEtc1[4] = {3, 2, 0, 1};

// Device-resident tables (built on the host, see build_tables below).
struct Tables {
  uint16_t inverse[64][256];  // [diff | inten<<1 | sel<<4][colour] = packed | err << 8
  uint16_t cfg_ends[2][33];   // exact configurations for colour 0 and 255
  uint16_t cfg[254][12];      // ... for colours 1..254, 0xFFFF-terminated
};

__device__ __forceinline__ int clamp255(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }
__device__ __forceinline__ int chan(uint32_t p, int k) { return (int)((p >> (8 * k)) & 0xFFu); }

typedef short s16x2_t __attribute__((ext_vector_type(2)));

// Red and green of texel p as two 16-bit lanes.
__device__ __forceinline__ uint32_t rg_of(uint32_t p) {
  return __builtin_amdgcn_perm(0u, p, 0x0c010c00u);
}

// Squared RGB distance of texel p to the colour (rg: red and green lanes,
// b): the red/green differences as 16-bit lanes in one two-way dot product
// (all terms below 2^17, so the sums are exact).
__device__ __forceinline__ uint32_t dist2(uint32_t p, uint32_t rg, int b) {
  const s16x2_t d = __builtin_bit_cast(s16x2_t, rg_of(p)) - __builtin_bit_cast(s16x2_t, rg);
  const int db = chan(p, 2) - b;
  return (uint32_t)__builtin_amdgcn_sdot2(d, d, db * db, false);
}

struct Solution {
  int r, g, b;      // unscaled block colour
  int inten;
  uint32_t sel;     // 8 selector indices, 2 bits each
  uint64_t err;
  bool valid;
};

struct Opt {
  uint32_t px[8];
  int luma_sorted[8];
  int idx_sorted[8];
  uint32_t px_sorted[8];  // px[idx_sorted[k]]
  float avg[3];
  int limit;
  int br, bg, bb;
  bool color4;
  bool constrain;
  int base_r, base_g, base_b;
};

// etc1_optimizer::init for 8 pixels (low quality: sorted luma needed).
__device__ void opt_init(Opt &o, const uint32_t sub[8], bool color4) {
  o.limit = color4 ? 15 : 31;
  o.color4 = color4;
  float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
  int luma[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    o.px[i] = sub[i];
    s0 += (float)chan(sub[i], 0);
    s1 += (float)chan(sub[i], 1);
    s2 += (float)chan(sub[i], 2);
    luma[i] = chan(sub[i], 0) + chan(sub[i], 1) + chan(sub[i], 2);
  }
  const float inv = 1.0f / 8.0f;
  o.avg[0] = s0 * inv;
  o.avg[1] = s1 * inv;
  o.avg[2] = s2 * inv;
  const float lim = (float)o.limit;
  o.br = min(max((int)(uint32_t)(o.avg[0] * lim / 255.0f + 0.5f), 0), o.limit);
  o.bg = min(max((int)(uint32_t)(o.avg[1] * lim / 255.0f + 0.5f), 0), o.limit);
  o.bb = min(max((int)(uint32_t)(o.avg[2] * lim / 255.0f + 0.5f), 0), o.limit);
  // stable sort of the 8 lumas (indirect_radix_sort is an LSD radix sort)
#pragma unroll
  for (int i = 0; i < 8; i++) {
    int rank = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) rank += (luma[j] < luma[i]) || (luma[j] == luma[i] && j < i);
    // scatter through a select chain (rank is 0..7)
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (k == rank) {
        o.luma_sorted[k] = luma[i];
        o.idx_sorted[k] = i;
        o.px_sorted[k] = sub[i];
      }
  }
}

__device__ __forceinline__ void scaled(const Opt &o, int r, int g, int b, int &sr, int &sg,
                                       int &sb) {
  if (o.color4) {
    sr = r | (r << 4);
    sg = g | (g << 4);
    sb = b | (b << 4);
  } else {
    sr = (r >> 2) | (r << 3);
    sg = (g >> 2) | (g << 3);
    sb = (b >> 2) | (b << 3);
  }
}

// evaluate_solution_fast; updates best and returns true on improvement.
__device__ bool evaluate_fast(const Opt &o, int cr, int cg, int cb, Solution &best) {
  if (o.constrain) {
    const int dr = cr - o.base_r, dg = cg - o.base_g, db = cb - o.base_b;
    const int mn = min(min(dr, dg), db), mx = max(max(dr, dg), db);
    if (mn < -4 || mx > 3) return false;
  }
  int sr, sg, sb;
  scaled(o, cr, cg, cb, sr, sg, sb);
  uint64_t terr = ~0ull;
  int tinten = 0;
  uint32_t tsel = 0;
  bool tvalid = false;
  const int lmax = o.luma_sorted[7], lmin = o.luma_sorted[0];
  for (int it = 7; it >= 0; --it) {
    // the four candidate colours of intensity row it (no private arrays:
    // a per-lane index into one would live in scratch memory)
    const int y0 = kInten[it][0], y1 = kInten[it][1], y2 = kInten[it][2], y3 = kInten[it][3];
    const int c0r = clamp255(sr + y0), c0g = clamp255(sg + y0), c0b = clamp255(sb + y0);
    const int c1r = clamp255(sr + y1), c1g = clamp255(sg + y1), c1b = clamp255(sb + y1);
    const int c2r = clamp255(sr + y2), c2g = clamp255(sg + y2), c2b = clamp255(sb + y2);
    const int c3r = clamp255(sr + y3), c3g = clamp255(sg + y3), c3b = clamp255(sb + y3);
    const int bi0 = c0r + c0g + c0b, bi1 = c1r + c1g + c1b, bi2 = c2r + c2g + c2b,
              bi3 = c3r + c3g + c3b;
    const int mid0 = bi0 + bi1, mid1 = bi1 + bi2, mid2 = bi2 + bi3;
    const uint32_t c0rg = (uint32_t)c0r | ((uint32_t)c0g << 16);
    const uint32_t c1rg = (uint32_t)c1r | ((uint32_t)c1g << 16);
    const uint32_t c2rg = (uint32_t)c2r | ((uint32_t)c2g << 16);
    const uint32_t c3rg = (uint32_t)c3r | ((uint32_t)c3g << 16);
    // a row's error is at most 8 * 3 * 255^2 < 2^21: 32-bit sums are exact
    uint32_t total = 0;
    uint32_t sel = 0;
    if (lmax * 2 < mid0) {
      if (bi0 > lmax && (uint64_t)(bi0 - lmax) >= terr) continue;
#pragma unroll
      for (int c = 0; c < 8; c++) total += dist2(o.px[c], c0rg, c0b);
      sel = 0;
    } else if (lmin * 2 >= mid2) {
      if (lmin > bi3 && (uint64_t)(lmin - bi3) >= terr) continue;
#pragma unroll
      for (int c = 0; c < 8; c++) total += dist2(o.px[c], c3rg, c3b);
      sel = 0xFFFFu;
    } else {
      // the reference walks cur up the midpoints as the sorted lumas rise;
      // the midpoints do not decrease (the row's offsets increase), so the
      // walk ends at the number of midpoints at or below 2 y: 3 if the last
      // is, else 2 if the middle one is, ...
#pragma unroll
      for (int c = 0; c < 8; c++) {
        const int y2x = o.luma_sorted[c] * 2;
        const bool g0 = y2x >= mid0, g1 = y2x >= mid1, g2 = y2x >= mid2;
        const uint32_t cur = g2 ? 3u : (g1 ? 2u : (g0 ? 1u : 0u));
        const int pi = o.idx_sorted[c];
        const uint32_t p = o.px_sorted[c];
        const uint32_t erg = g2 ? c3rg : (g1 ? c2rg : (g0 ? c1rg : c0rg));
        const int eb = g2 ? c3b : (g1 ? c2b : (g0 ? c1b : c0b));
        total += dist2(p, erg, eb);
        sel |= cur << (2 * pi);
      }
    }
    if (total < terr) {
      terr = total;
      tinten = it;
      tsel = sel;
      tvalid = true;
      if (!total) break;
    }
  }
  if (tvalid && terr < best.err) {
    best.err = terr;
    best.inten = tinten;
    best.sel = tsel;
    best.r = cr;
    best.g = cg;
    best.b = cb;
    best.valid = true;
    return true;
  }
  return false;
}

// etc1_optimizer::compute with scan deltas {0} and two refinement trials.
__device__ bool opt_compute(const Opt &o, Solution &best) {
  best.valid = false;
  best.err = ~0ull;
  best.r = best.g = best.b = 0;
  best.inten = 0;
  best.sel = 0;
  const int mbr = o.br, mbg = o.bg, mbb = o.bb;
  if (evaluate_fast(o, mbr, mbg, mbb, best)) {
    for (int trial = 0; trial < 2; trial++) {
      int sr, sg, sb;
      scaled(o, best.r, best.g, best.b, sr, sg, sb);
      int dsr = 0, dsg = 0, dsb = 0;
#pragma unroll
      for (int r = 0; r < 8; r++) {
        const int yd = kInten[best.inten][(best.sel >> (2 * r)) & 3];
        dsr += clamp255(sr + yd) - sr;
        dsg += clamp255(sg + yd) - sg;
        dsb += clamp255(sb + yd) - sb;
      }
      if (!dsr && !dsg && !dsb) break;
      const float adr = (float)dsr / 8.0f, adg = (float)dsg / 8.0f, adb = (float)dsb / 8.0f;
      const float lim = (float)o.limit;
      const int br1 = min(max((int)(uint32_t)((o.avg[0] - adr) * lim / 255.0f + 0.5f), 0), o.limit);
      const int bg1 = min(max((int)(uint32_t)((o.avg[1] - adg) * lim / 255.0f + 0.5f), 0), o.limit);
      const int bb1 = min(max((int)(uint32_t)((o.avg[2] - adb) * lim / 255.0f + 0.5f), 0), o.limit);
      if (mbr == br1 && mbg == bg1 && mbb == bb1) break;
      if (br1 == best.r && bg1 == best.g && bb1 == best.b) break;
      if (o.br == br1 && o.bg == bg1 && o.bb == bb1) break;
      if (!evaluate_fast(o, br1, bg1, bb1, best)) break;
    }
  }
  return best.valid;
}

// pack_etc1_block_solid_color
__device__ uint2 pack_solid(const Tables *t, uint32_t pixel) {
  const int next[4] = {1, 2, 0, 1};
  int col[3] = {chan(pixel, 0), chan(pixel, 1), chan(pixel, 2)};
  uint32_t best_err = 0xFFFFFFFFu, best_x = 0, best_c1 = 0, best_c2 = 0;
  int best_i = 0;
  bool done = false;
  for (int i = 0; i < 3 && !done; i++) {
    const int c1 = col[next[i]], c2 = col[next[i + 1]];
    for (int delta = -1; delta <= 1 && !done; delta++) {
      const int cpd = clamp255(col[i] + delta);
      const uint16_t *tab = cpd == 0 ? t->cfg_ends[0] : (cpd == 255 ? t->cfg_ends[1] : t->cfg[cpd - 1]);
      int k = 0;
      do {
        const uint32_t x = tab[k++];
        const uint16_t *inv = t->inverse[x & 0xFF];
        const uint32_t p1 = inv[c1], p2 = inv[c2];
        const uint32_t e1 = p1 >> 8, e2 = p2 >> 8;
        const uint32_t err = (uint32_t)((cpd - col[i]) * (cpd - col[i])) + e1 * e1 + e2 * e2;
        if (err < best_err) {
          best_err = err;
          best_x = x;
          best_c1 = p1 & 0xFF;
          best_c2 = p2 & 0xFF;
          best_i = i;
          if (!best_err) {
            done = true;
            break;
          }
        }
      } while (tab[k] != 0xFFFF);
    }
  }
  const uint32_t diff = best_x & 1, inten = (best_x >> 1) & 7;
  uint8_t bytes[8];
  bytes[3] = (uint8_t)(((inten | (inten << 3)) << 2) | (diff << 1));
  const uint32_t es = kSelToEtc1[(best_x >> 4) & 3];
  bytes[4] = bytes[5] = (es & 2) ? 0xFF : 0;
  bytes[6] = bytes[7] = (es & 1) ? 0xFF : 0;
  const uint32_t c0 = (best_x >> 8) & 255;
  uint32_t v0, v1, v2;
  if (diff) {
    v0 = c0 << 3;
    v1 = best_c1 << 3;
    v2 = best_c2 << 3;
  } else {
    v0 = c0 | (c0 << 4);
    v1 = best_c1 | (best_c1 << 4);
    v2 = best_c2 | (best_c2 << 4);
  }
  // bytes[best_i] = v0, bytes[next[best_i]] = v1, bytes[next[best_i + 1]] = v2
  uint32_t b012[3];
  b012[best_i] = v0;
  b012[next[best_i]] = v1;
  b012[next[best_i + 1]] = v2;
  bytes[0] = (uint8_t)b012[0];
  bytes[1] = (uint8_t)b012[1];
  bytes[2] = (uint8_t)b012[2];
  uint2 r;
  r.x = bytes[0] | (bytes[1] << 8) | (bytes[2] << 16) | ((uint32_t)bytes[3] << 24);
  r.y = bytes[4] | (bytes[5] << 8) | (bytes[6] << 16) | ((uint32_t)bytes[7] << 24);
  return r;
}

__device__ __forceinline__ void push_sel(uint32_t &s0, uint32_t &s1, uint32_t sel_idx) {
  const uint32_t b = kSelToEtc1[sel_idx & 3];
  s0 = (s0 << 1) | (b & 1);
  s1 = (s1 << 1) | (b >> 1);
}

// One sub-block's fit (the body of pack_etc1_block's subblock loop,
// rg_etc1.cpp:2321-2352): its 8 pixels in the reference's order, gathered
// with static indices and selects, then etc1_optimizer.
__device__ __forceinline__ bool fit_subblock(const uint32_t px[16], int flip, int sb, bool color4,
                                             bool constrain, int base_r, int base_g, int base_b,
                                             Solution &res) {
  uint32_t sub[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    // flipped: rows 2 sb, 2 sb + 1; otherwise columns 2 sb, 2 sb + 1
    const uint32_t f = sb ? px[8 + k] : px[k];
    const int q = (k >> 2) + 4 * (k & 3);
    const uint32_t u = sb ? px[2 + q] : px[q];
    sub[k] = flip ? f : u;
  }
  Opt o;
  opt_init(o, sub, color4);
  o.constrain = constrain;
  o.base_r = base_r;
  o.base_g = base_g;
  o.base_b = base_b;
  return opt_compute(o, res);
}

__device__ uint2 pack_block(const Tables *t, const uint32_t px[16]) {
  bool solid = true;
#pragma unroll
  for (int r = 1; r < 16; r++) solid = solid && px[r] == px[0];
  if (solid) return pack_solid(t, px[0]);

  uint64_t best_error = ~0ull;
  int best_flip = 0, best_c4 = 0;
  Solution best0, best1;
  best0.valid = best1.valid = false;
  best0.r = best0.g = best0.b = best0.inten = 0;
  best0.sel = 0;
  best0.err = 0;
  best1 = best0;
  for (int flip = 0; flip < 2; flip++) {
    for (int c4 = 0; c4 < 2; c4++) {
      Solution r0, r1;
      if (!fit_subblock(px, flip, 0, c4 != 0, false, 0, 0, 0, r0)) continue;
      uint64_t trial_error = r0.err;
      if (trial_error >= best_error) continue;
      // the 555 mode's second colour is a delta of the first: constrained
      if (!fit_subblock(px, flip, 1, c4 != 0, !c4, r0.r, r0.g, r0.b, r1)) continue;
      trial_error += r1.err;
      if (trial_error >= best_error) continue;
      best_error = trial_error;
      best0 = r0;
      best1 = r1;
      best_flip = flip;
      best_c4 = c4;
    }
  }
  Solution best_res[2] = {best0, best1};
  uint8_t b0, b1, b2;
  if (best_c4) {
    b0 = (uint8_t)(best_res[1].r | (best_res[0].r << 4));
    b1 = (uint8_t)(best_res[1].g | (best_res[0].g << 4));
    b2 = (uint8_t)(best_res[1].b | (best_res[0].b << 4));
  } else {
    int dr = best_res[1].r - best_res[0].r;
    int dg = best_res[1].g - best_res[0].g;
    int db = best_res[1].b - best_res[0].b;
    if (dr < 0) dr += 8;
    if (dg < 0) dg += 8;
    if (db < 0) db += 8;
    b0 = (uint8_t)((best_res[0].r << 3) | dr);
    b1 = (uint8_t)((best_res[0].g << 3) | dg);
    b2 = (uint8_t)((best_res[0].b << 3) | db);
  }
  const uint8_t b3 = (uint8_t)((best_res[1].inten << 2) | (best_res[0].inten << 5) |
                               ((~best_c4 & 1) << 1) | best_flip);
  uint32_t s0 = 0, s1 = 0;
  const uint32_t q0 = best_res[0].sel, q1 = best_res[1].sel;
  if (best_flip) {
    for (int x = 3; x >= 0; --x) {
      push_sel(s0, s1, q1 >> (2 * (4 + x)));
      push_sel(s0, s1, q1 >> (2 * x));
      push_sel(s0, s1, q0 >> (2 * (4 + x)));
      push_sel(s0, s1, q0 >> (2 * x));
    }
  } else {
    for (int sbk = 1; sbk >= 0; --sbk) {
      const uint32_t q = sbk ? q1 : q0;
      for (int i = 1; i >= 0; --i) {
        const int o = 4 * i;
        push_sel(s0, s1, q >> (2 * (o + 3)));
        push_sel(s0, s1, q >> (2 * (o + 2)));
        push_sel(s0, s1, q >> (2 * (o + 1)));
        push_sel(s0, s1, q >> (2 * (o + 0)));
      }
    }
  }
  uint2 r;
  r.x = b0 | (b1 << 8) | (b2 << 16) | ((uint32_t)b3 << 24);
  r.y = ((s1 >> 8) & 0xFF) | ((s1 & 0xFF) << 8) | (((s0 >> 8) & 0xFF) << 16) | ((s0 & 0xFF) << 24);
  return r;
}

__global__ void __launch_bounds__(256) encode_kernel(BlockRange r, const Tables *t) {
  const uint32_t rel = blockIdx.x * blockDim.x + threadIdx.x;
  if (rel >= r.num_blocks) return;
  const uint32_t b = r.first_block + rel;
  const uint32_t bx = b % r.blocks_wide, by = b / r.blocks_wide;
  const uint32_t *p = r.img + (size_t)(by * 4) * r.width + bx * 4;
  uint32_t px[16];
#pragma unroll
  for (int y = 0; y < 4; y++) {
    const uint4 row = *reinterpret_cast<const uint4 *>(p + (size_t)y * r.width);
    px[y * 4 + 0] = row.x;
    px[y * 4 + 1] = row.y;
    px[y * 4 + 2] = row.z;
    px[y * 4 + 3] = row.w;
  }
  reinterpret_cast<uint2 *>(r.out)[b] = pack_block(t, px);
}

// Batch form: persistent workgroups take chunks of the group's blocks from
// a device counter; each block finds its surface through block_surf.
// Workgroups of at most 512 threads (the autotuner's largest shape).
__global__ void __launch_bounds__(512) batch_kernel(const SurfDesc *surfs, const uint32_t *block_surf, uint32_t total,
                             uint32_t *counter, const Tables *t) {
  __shared__ uint32_t s_chunk;
  for (;;) {
    if (threadIdx.x == 0) s_chunk = atomicAdd(counter, blockDim.x);
    __syncthreads();
    const uint32_t chunk = s_chunk;
    __syncthreads();
    if (chunk >= total) break;
    const uint32_t g = chunk + threadIdx.x;
    if (g >= total) continue;
    const SurfDesc sd = surfs[block_surf[g]];
    const uint32_t b = g - sd.first;
    const uint32_t bx = b % sd.blocks_wide, by = b / sd.blocks_wide;
    const uint32_t *p = sd.img + (size_t)(by * 4) * sd.width + bx * 4;
    uint32_t px[16];
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const uint4 row = *reinterpret_cast<const uint4 *>(p + (size_t)y * sd.width);
      px[y * 4 + 0] = row.x;
      px[y * 4 + 1] = row.y;
      px[y * 4 + 2] = row.z;
      px[y * 4 + 3] = row.w;
    }
    reinterpret_cast<uint2 *>(sd.out)[b] = pack_block(t, px);
  }
}

}  // namespace etc

namespace {

const int kInt[8][4] = {{-8, -2, 2, 8},       {-17, -5, 5, 17},   {-29, -9, 9, 29},
                        {-42, -13, 13, 42},   {-60, -18, 18, 60}, {-80, -24, 24, 80},
                        {-106, -33, 33, 106}, {-183, -47, 47, 183}};

int decode_value(int diff, int inten, int sel, int packed) {
  int c = diff ? ((packed >> 2) | (packed << 3)) : (packed | (packed << 4));
  c += kInt[inten][sel];
  return c < 0 ? 0 : (c > 255 ? 255 : c);
}

// pack_etc1_block_init's inverse table (rg_etc1.cpp:1905-1949) and the exact
// configuration lists of g_color8_to_etc_block_config_* (rg_etc1.cpp:385-503):
// every (diff, intensity, selector, packed colour) that decodes to the colour,
// enumerated diff-major, intensity ascending, selector 3 down to 0, packed
// colour ascending.  Colour 0 lists selectors 0 and 1 with packed colour 0;
// colour 255 lists selectors 2 and 3 with the smallest packed colour that
// reaches 255, ordered by that packed colour.
void build_tables(etc::Tables &t) {
  for (int diff = 0; diff < 2; diff++) {
    const int limit = diff ? 32 : 16;
    for (int inten = 0; inten < 8; inten++)
      for (int sel = 0; sel < 4; sel++) {
        const int idx = diff + (inten << 1) + (sel << 4);
        for (int color = 0; color < 256; color++) {
          uint32_t best = 0xFFFFFFFFu, bp = 0;
          for (int p = 0; p < limit; p++) {
            const int v = decode_value(diff, inten, sel, p);
            const uint32_t e = (uint32_t)(v > color ? v - color : color - v);
            if (e < best) {
              best = e;
              bp = (uint32_t)p;
              if (!best) break;
            }
          }
          t.inverse[idx][color] = (uint16_t)(bp | (best << 8));
        }
      }
  }
  for (int color = 1; color < 255; color++) {
    int n = 0;
    for (int diff = 0; diff < 2; diff++)
      for (int inten = 0; inten < 8; inten++)
        for (int sel = 3; sel >= 0; sel--)
          for (int p = 0; p < (diff ? 32 : 16); p++)
            if (decode_value(diff, inten, sel, p) == color && n < 11)
              t.cfg[color - 1][n++] = (uint16_t)(diff | (inten << 1) | (sel << 4) | (p << 8));
    for (; n < 12; n++) t.cfg[color - 1][n] = 0xFFFF;
  }
  int n = 0;
  for (int diff = 0; diff < 2; diff++)
    for (int inten = 0; inten < 8; inten++)
      for (int sel = 0; sel < 2; sel++) t.cfg_ends[0][n++] = (uint16_t)(diff | (inten << 1) | (sel << 4));
  t.cfg_ends[0][n] = 0xFFFF;
  n = 0;
  for (int diff = 0; diff < 2; diff++)
    for (int inten = 0; inten < 8; inten++) {
      int ps[2];
      for (int k = 0; k < 2; k++) {
        const int sel = 2 + k;
        ps[k] = 0;
        for (int p = 0; p < (diff ? 32 : 16); p++)
          if (decode_value(diff, inten, sel, p) == 255) {
            ps[k] = p;
            break;
          }
      }
      const int first = (ps[1] < ps[0]) ? 1 : 0;  // by packed colour, then selector
      for (int k = 0; k < 2; k++) {
        const int kk = k == 0 ? first : 1 - first;
        t.cfg_ends[1][n++] = (uint16_t)(diff | (inten << 1) | ((2 + kk) << 4) | (ps[kk] << 8));
      }
    }
  t.cfg_ends[1][n] = 0xFFFF;
}

etc::Tables *g_tables[64];
int g_etc_cus[64];

}  // namespace

hipError_t etc1_prepare() {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  if (!g_tables[dev]) {
    etc::Tables *h = new etc::Tables();
    build_tables(*h);
    etc::Tables *d = nullptr;
    e = hipMalloc(&d, sizeof(etc::Tables));
    if (e == hipSuccess) e = hipMemcpy(d, h, sizeof(etc::Tables), hipMemcpyHostToDevice);
    delete h;
    if (e != hipSuccess) return e;
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    g_etc_cus[dev] = cus > 0 ? cus : 1;
    g_tables[dev] = d;
  }
  return hipSuccess;
}

hipError_t etc1_encode_batch(const SurfDesc *surfs, const uint32_t *block_surf, uint32_t total,
                             uint32_t *counter, const LaunchTune &tune, hipStream_t stream) {
  if (total == 0) return hipSuccess;
  hipError_t e = etc1_prepare();
  if (e != hipSuccess) return e;
  int dev = 0;
  (void)hipGetDevice(&dev);
  const uint32_t wg = tune.simple_wg;
  uint32_t grid = (uint32_t)g_etc_cus[dev] * (2048u / wg);
  if (grid > (total + wg - 1) / wg) grid = (total + wg - 1) / wg;
  const int t = timing_begin(stream);
  hipLaunchKernelGGL(etc::batch_kernel, dim3(grid), dim3(wg), 0, stream, surfs, block_surf,
                     total, counter, (const etc::Tables *)g_tables[dev]);
  timing_end(t, kTagEtc1, stream);
  return hipGetLastError();
}

hipError_t etc1_encode(const BlockRange &r, hipStream_t stream) {
  if (r.num_blocks == 0) return hipSuccess;
  int dev = 0;
  hipError_t e = etc1_prepare();
  if (e != hipSuccess) return e;
  (void)hipGetDevice(&dev);
  const uint32_t grid = (r.num_blocks + 255) / 256;
  const int t = timing_begin(stream);
  hipLaunchKernelGGL(etc::encode_kernel, dim3(grid), dim3(256), 0, stream, r, g_tables[dev]);
  timing_end(t, kTagEtc1, stream);
  return hipGetLastError();
}

}  // namespace fh
