// DXT1 / DXT5 (BC1 / BC3) block encoder for gfx950.
//
// FasTC encodes DXT with stb_dxt v1.06 in dithering mode
// (DXTEncoder/src/Compressor.cpp:47-95: stb_compress_dxt_block(out, block,
// alpha, STB_DXT_DITHER)).  This kernel reproduces that encoder bit for bit:
// Floyd-Steinberg dither to 565 (stb_dxt.h:159-183), integer covariance and
// a 4-step float power iteration (:283-385), dithered index matching
// (:186-280), one least-squares refinement (:398-474) and the DXT5 alpha
// block (:551-601).
//
// One lane per 4x4 block; each lane loads its four 16-byte rows with 128-bit
// loads, so a wavefront over 64 horizontally adjacent blocks reads four
// contiguous 1 KiB runs.  The encode is a few hundred integer operations per
// block, so the kernel is bound by HBM traffic (4 B/texel in, 0.5 or 1 B/texel
// out).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.h"

namespace fh {
namespace dxt {

struct Tables {
  uint8_t expand5[32];
  uint8_t expand6[64];
  uint8_t omatch5[256][2];
  uint8_t omatch6[256][2];
  // stb__QuantRBTab / QuantGTab over the dither's whole index range
  // [-256, 527], offset by 256: quant[i + 256] = the 5-6-bit quantisation of
  // clamp(i - 8, 0, 255), expanded back to 8 bits
  uint8_t quant_rb[784];
  uint8_t quant_g[784];
};

__constant__ Tables kT;

__device__ __forceinline__ int mul8bit(int a, int b) {
  const int t = a * b + 128;
  return (t + (t >> 8)) >> 8;
}

__device__ __forceinline__ uint32_t as16bit(int r, int g, int b) {
  return (uint32_t)((mul8bit(r, 31) << 11) + (mul8bit(g, 63) << 5) + mul8bit(b, 31));
}

__device__ __forceinline__ int ch(uint32_t p, int k) { return (int)((p >> (8 * k)) & 0xFFu); }

typedef short short2_t __attribute__((ext_vector_type(2)));

// r * vr + g * vg + b * vb for texel p, exactly: red and green as a pair of
// 16-bit lanes in one two-way dot product (|v| < 2^15)
__device__ __forceinline__ int dot3(uint32_t p, int vr, int vg, int vb) {
  const uint32_t rg = __builtin_amdgcn_perm(0u, p, 0x0c010c00u);  // r, 0, g, 0
  const uint32_t w = ((uint32_t)vr & 0xFFFFu) | ((uint32_t)vg << 16);
  return __builtin_amdgcn_sdot2(__builtin_bit_cast(short2_t, rg), __builtin_bit_cast(short2_t, w),
                                ch(p, 2) * vb, false);
}

// stb__EvalColors: color[0] = max16, color[1] = min16, color[2..3] = thirds.
__device__ __forceinline__ void eval_colors(int col[4][3], uint32_t c0, uint32_t c1) {
  col[0][0] = kT.expand5[(c0 >> 11) & 31];
  col[0][1] = kT.expand6[(c0 >> 5) & 63];
  col[0][2] = kT.expand5[c0 & 31];
  col[1][0] = kT.expand5[(c1 >> 11) & 31];
  col[1][1] = kT.expand6[(c1 >> 5) & 63];
  col[1][2] = kT.expand5[c1 & 31];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    col[2][k] = (2 * col[0][k] + col[1][k]) / 3;
    col[3][k] = (2 * col[1][k] + col[0][k]) / 3;
  }
}

// The kernels copy the two tables into LDS (q: red/blue then green, 784
// bytes each), where the lookups of the dither's serial chain are short.
__device__ __forceinline__ int qtab(const uint8_t *q, int ch_, int idx) {
  return q[(ch_ == 1 ? 784 : 0) + idx + 256];
}

__device__ __forceinline__ void load_quant(uint8_t *q) {
  for (int i = (int)threadIdx.x; i < 784; i += (int)blockDim.x) {
    q[i] = kT.quant_rb[i];
    q[784 + i] = kT.quant_g[i];
  }
  __syncthreads();
}

// stb__DitherBlock
__device__ void dither_block(const uint32_t blk[16], uint32_t dst[16], const uint8_t *q) {
#pragma unroll
  for (int i = 0; i < 16; i++) dst[i] = 0;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    int e1[4] = {0, 0, 0, 0}, e2[4] = {0, 0, 0, 0};
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const int b0 = ch(blk[y * 4 + 0], c), b1 = ch(blk[y * 4 + 1], c);
      const int b2 = ch(blk[y * 4 + 2], c), b3 = ch(blk[y * 4 + 3], c);
      // ep1 = e1 (current row), ep2 = e2 (previous row)
      const int d0 = qtab(q, c, b0 + ((3 * e2[1] + 5 * e2[0]) >> 4) + 8);
      e1[0] = b0 - d0;
      const int d1 = qtab(q, c, b1 + ((7 * e1[0] + 3 * e2[2] + 5 * e2[1] + e2[0]) >> 4) + 8);
      e1[1] = b1 - d1;
      const int d2 = qtab(q, c, b2 + ((7 * e1[1] + 3 * e2[3] + 5 * e2[2] + e2[1]) >> 4) + 8);
      e1[2] = b2 - d2;
      const int d3 = qtab(q, c, b3 + ((7 * e1[2] + 5 * e2[3] + e2[2]) >> 4) + 8);
      e1[3] = b3 - d3;
      dst[y * 4 + 0] |= (uint32_t)d0 << (8 * c);
      dst[y * 4 + 1] |= (uint32_t)d1 << (8 * c);
      dst[y * 4 + 2] |= (uint32_t)d2 << (8 * c);
      dst[y * 4 + 3] |= (uint32_t)d3 << (8 * c);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int t = e1[k];
        e1[k] = e2[k];
        e2[k] = t;
      }
    }
  }
}

// One dithered index choice of stb__MatchColorsBlock: dot is the pixel's
// projection (times 16, error feedback added).  The index is 1 or 3 below the
// half-way point and 2 or 0 above it; it goes to mask at bit pos.  Returns the
// pixel's projection minus that of its palette entry (s0..s3: the entries'
// projections).
__device__ __forceinline__ int match_step(int dp, int dot, int half, int c0p, int c3p, int s0,
                                          int s1, int s2, int s3, uint32_t &mask, int pos) {
  const bool below = dot < c0p, above = dot < c3p;
  const int t_lo = below ? s1 : s3, t_hi = above ? s2 : s0;
  const uint32_t i_lo = below ? 1u : 3u, i_hi = above ? 2u : 0u;
  const bool lo = dot < half;
  mask |= (lo ? i_lo : i_hi) << pos;
  return dp - (lo ? t_lo : t_hi);
}

// stb__MatchColorsBlock with dithering.
__device__ uint32_t match_colors(const uint32_t blk[16], const int col[4][3]) {
  const int dirr = col[0][0] - col[1][0];
  const int dirg = col[0][1] - col[1][1];
  const int dirb = col[0][2] - col[1][2];
  int dots[16], stops[4];
#pragma unroll
  for (int i = 0; i < 16; i++) dots[i] = dot3(blk[i], dirr, dirg, dirb);
#pragma unroll
  for (int i = 0; i < 4; i++) stops[i] = col[i][0] * dirr + col[i][1] * dirg + col[i][2] * dirb;
  const int c0p = ((stops[1] + stops[3]) >> 1) << 4;
  const int half = ((stops[3] + stops[2]) >> 1) << 4;
  const int c3p = ((stops[2] + stops[0]) >> 1) << 4;
  int e1[4] = {0, 0, 0, 0}, e2[4] = {0, 0, 0, 0};
  uint32_t mask = 0;
#pragma unroll
  for (int y = 0; y < 4; y++) {
    const int *dp = &dots[y * 4];
    const int p = y * 8;
    e1[0] = match_step(dp[0], (dp[0] << 4) + (3 * e2[1] + 5 * e2[0]), half, c0p, c3p, stops[0],
                       stops[1], stops[2], stops[3], mask, p);
    e1[1] = match_step(dp[1], (dp[1] << 4) + (7 * e1[0] + 3 * e2[2] + 5 * e2[1] + e2[0]), half,
                       c0p, c3p, stops[0], stops[1], stops[2], stops[3], mask, p + 2);
    e1[2] = match_step(dp[2], (dp[2] << 4) + (7 * e1[1] + 3 * e2[3] + 5 * e2[2] + e2[1]), half,
                       c0p, c3p, stops[0], stops[1], stops[2], stops[3], mask, p + 4);
    e1[3] = match_step(dp[3], (dp[3] << 4) + (7 * e1[2] + 5 * e2[3] + e2[2]), half, c0p, c3p,
                       stops[0], stops[1], stops[2], stops[3], mask, p + 6);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int t = e1[k];
      e1[k] = e2[k];
      e2[k] = t;
    }
  }
  return mask;
}

// stb__OptimizeColorsBlock
__device__ void optimize_colors(const uint32_t blk[16], uint32_t &max16, uint32_t &min16) {
  int mu[3], mn[3], mx[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    int muv = ch(blk[0], c), minv = muv, maxv = muv;
#pragma unroll
    for (int i = 1; i < 16; i++) {
      const int v = ch(blk[i], c);
      muv += v;
      // stb_dxt's "if (v < min) ... else if (v > max)" with min <= max
      // throughout is the plain running minimum and maximum
      minv = min(minv, v);
      maxv = max(maxv, v);
    }
    mu[c] = (muv + 8) >> 4;
    mn[c] = minv;
    mx[c] = maxv;
  }
  // sum (x - mx)(y - my) = sum xy - my sum x - mx sum y + 16 mx my, exactly
  // in integers: the products are of bytes
  int sp[6] = {0, 0, 0, 0, 0, 0}, sx[3] = {0, 0, 0};
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int r = ch(blk[i], 0), g = ch(blk[i], 1), b = ch(blk[i], 2);
    sp[0] += r * r;
    sp[1] += r * g;
    sp[2] += r * b;
    sp[3] += g * g;
    sp[4] += g * b;
    sp[5] += b * b;
    sx[0] += r;
    sx[1] += g;
    sx[2] += b;
  }
  int cov[6];
  {
    const int pr[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const int x = pr[k][0], y = pr[k][1];
      cov[k] = sp[k] - mu[y] * sx[x] - mu[x] * sx[y] + 16 * mu[x] * mu[y];
    }
  }
  // cov / 255.0f, correctly rounded: |cov| < 2^24, and cov / 255 is an
  // integer or not a dyadic rational, at least 2^-33 (relative) from any
  // float rounding boundary, so the double product (within 2^-52 of it)
  // rounds to the same float
  float covf[6];
#pragma unroll
  for (int i = 0; i < 6; i++) covf[i] = (float)((double)cov[i] * (1.0 / 255.0));
  float vfr = (float)(mx[0] - mn[0]);
  float vfg = (float)(mx[1] - mn[1]);
  float vfb = (float)(mx[2] - mn[2]);
#pragma unroll
  for (int it = 0; it < 4; it++) {
    const float r = vfr * covf[0] + vfg * covf[1] + vfb * covf[2];
    const float g = vfr * covf[1] + vfg * covf[3] + vfb * covf[4];
    const float b = vfr * covf[2] + vfg * covf[4] + vfb * covf[5];
    vfr = r;
    vfg = g;
    vfb = b;
  }
  double magn = fabs((double)vfr);
  if (fabs((double)vfg) > magn) magn = fabs((double)vfg);
  if (fabs((double)vfb) > magn) magn = fabs((double)vfb);
  int v_r, v_g, v_b;
  if (magn < 4.0) {
    v_r = 299;
    v_g = 587;
    v_b = 114;
  } else {
    magn = 512.0 / magn;
    v_r = (int)((double)vfr * magn);
    v_g = (int)((double)vfg * magn);
    v_b = (int)((double)vfb * magn);
  }
  // |v| <= 587 already: the clamp only tells the compiler that 24-bit
  // multiplies are exact
  v_r = min(max(v_r, -1024), 1024);
  v_g = min(max(v_g, -1024), 1024);
  v_b = min(max(v_b, -1024), 1024);
  int mind = 0x7fffffff, maxd = -0x7fffffff;
  uint32_t minp = blk[0], maxp = blk[0];
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int dot = dot3(blk[i], v_r, v_g, v_b);
    if (dot < mind) {
      mind = dot;
      minp = blk[i];
    }
    if (dot > maxd) {
      maxd = dot;
      maxp = blk[i];
    }
  }
  max16 = as16bit(ch(maxp, 0), ch(maxp, 1), ch(maxp, 2));
  min16 = as16bit(ch(minp, 0), ch(minp, 1), ch(minp, 2));
}

__device__ __forceinline__ int sclamp(float y, int p0, int p1) {
  const int x = (int)y;
  return x < p0 ? p0 : (x > p1 ? p1 : x);
}

// stb__RefineBlock; returns whether the endpoints changed.
__device__ bool refine_block(const uint32_t blk[16], uint32_t &max16, uint32_t &min16,
                             uint32_t mask) {
  const uint32_t old_min = min16, old_max = max16;
  uint32_t nmin, nmax;
  if ((mask ^ (mask << 2)) < 4) {
    int r = 8, g = 8, b = 8;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      r += ch(blk[i], 0);
      g += ch(blk[i], 1);
      b += ch(blk[i], 2);
    }
    r >>= 4;
    g >>= 4;
    b >>= 4;
    nmax = ((uint32_t)kT.omatch5[r][0] << 11) | ((uint32_t)kT.omatch6[g][0] << 5) | kT.omatch5[b][0];
    nmin = ((uint32_t)kT.omatch5[r][1] << 11) | ((uint32_t)kT.omatch6[g][1] << 5) | kT.omatch5[b][1];
  } else {
    // index k's weights are w1 = 3, 0, 2, 1 (k = 0..3) and w2 = 3 - w1:
    // the sums of w1*w1, w2*w2 and w1*w2 follow from the index counts
    // (stb_dxt accumulates them as the bytes of one word, which never carry)
    const uint32_t lo = mask & 0x55555555u, hi = (mask >> 1) & 0x55555555u;
    const int n1 = __popc(lo & ~hi), n2 = __popc(hi & ~lo), n3 = __popc(lo & hi);
    const int n0 = 16 - n1 - n2 - n3;
    const int xx = 9 * n0 + 4 * n2 + n3;
    const int yy = 9 * n1 + n2 + 4 * n3;
    const int xy = 2 * (n2 + n3);
    // red and blue in the two 16-bit halves of one word: each half's sum
    // stays below 3 * 16 * 255 < 2^16, and r | b << 16 < 2^24
    uint32_t a1rb = 0, a2rb = 0, a1g = 0, a2g = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const uint32_t w1 = (0x1203u >> (((mask >> (2 * i)) & 3u) * 4u)) & 15u;
      const uint32_t rb = blk[i] & 0x00FF00FFu, g = (blk[i] >> 8) & 0xFFu;
      a1rb += w1 * rb;
      a1g += w1 * g;
      a2rb += rb;
      a2g += g;
    }
    const int a1r = (int)(a1rb & 0xFFFFu), a1b = (int)(a1rb >> 16), a1gi = (int)a1g;
    const int a2r = 3 * (int)(a2rb & 0xFFFFu) - a1r;
    const int a2gi = 3 * (int)a2g - a1gi;
    const int a2b = 3 * (int)(a2rb >> 16) - a1b;
    // 3.0f * 31.0f / 255.0f is evaluated left to right in float
    const float frb = ((3.0f * 31.0f) / 255.0f) / (float)(xx * yy - xy * xy);
    const float fg = (frb * 63.0f) / 31.0f;
    nmax = (uint32_t)sclamp((float)(a1r * yy - a2r * xy) * frb + 0.5f, 0, 31) << 11;
    nmax |= (uint32_t)sclamp((float)(a1gi * yy - a2gi * xy) * fg + 0.5f, 0, 63) << 5;
    nmax |= (uint32_t)sclamp((float)(a1b * yy - a2b * xy) * frb + 0.5f, 0, 31);
    nmin = (uint32_t)sclamp((float)(a2r * xx - a1r * xy) * frb + 0.5f, 0, 31) << 11;
    nmin |= (uint32_t)sclamp((float)(a2gi * xx - a1gi * xy) * fg + 0.5f, 0, 63) << 5;
    nmin |= (uint32_t)sclamp((float)(a2b * xx - a1b * xy) * frb + 0.5f, 0, 31);
  }
  min16 = nmin;
  max16 = nmax;
  return old_min != nmin || old_max != nmax;
}

// stb__CompressColorBlock in STB_DXT_DITHER mode (one refinement pass).
__device__ uint2 compress_color(const uint32_t blk[16], const uint8_t *q) {
  uint32_t mask, max16, min16;
  bool constant = true;
#pragma unroll
  for (int i = 1; i < 16; i++) constant = constant && blk[i] == blk[0];
  if (constant) {
    const int r = ch(blk[0], 0), g = ch(blk[0], 1), b = ch(blk[0], 2);
    mask = 0xaaaaaaaau;
    max16 = ((uint32_t)kT.omatch5[r][0] << 11) | ((uint32_t)kT.omatch6[g][0] << 5) | kT.omatch5[b][0];
    min16 = ((uint32_t)kT.omatch5[r][1] << 11) | ((uint32_t)kT.omatch6[g][1] << 5) | kT.omatch5[b][1];
  } else {
    uint32_t dblk[16];
    dither_block(blk, dblk, q);
    optimize_colors(dblk, max16, min16);
    int col[4][3];
    if (max16 != min16) {
      eval_colors(col, max16, min16);
      mask = match_colors(blk, col);
    } else {
      mask = 0;
    }
    const uint32_t last = mask;
    if (refine_block(dblk, max16, min16, mask)) {
      if (max16 != min16) {
        eval_colors(col, max16, min16);
        mask = match_colors(blk, col);
      } else {
        mask = 0;
      }
    }
    (void)last;
  }
  if (max16 < min16) {
    const uint32_t t = min16;
    min16 = max16;
    max16 = t;
    mask ^= 0x55555555u;
  }
  return make_uint2(max16 | (min16 << 16), mask);
}

// stb__CompressAlphaBlock
__device__ uint2 compress_alpha(const uint32_t blk[16]) {
  int mn = (int)(blk[0] >> 24), mx = mn;
#pragma unroll
  for (int i = 1; i < 16; i++) {
    const int a = (int)(blk[i] >> 24);
    mn = min(mn, a);  // stb_dxt's if / else-if, with mn <= mx throughout
    mx = max(mx, a);
  }
  const int dist = mx - mn;
  int bias = dist < 8 ? dist - 1 : dist / 2 + 2;
  bias -= mn * 7;
  // stb_dxt finds the index by binary long division of a = 7 v + bias, which
  // lies in [0, 8 dist) when dist > 0: q = floor(a / dist) = (a * m) >> 19
  // with m = ceil(2^19 / dist), exact for a < 2^11.  When dist = 0, a = -1
  // and stb_dxt's steps give q = 0, as m = 0 does.
  const uint32_t m = dist ? ((1u << 19) + (uint32_t)dist - 1u) / (uint32_t)dist : 0u;
  uint64_t bits = 0;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const uint32_t a = (uint32_t)((int)(blk[i] >> 24) * 7 + bias) & 0x7FFu;
    const uint32_t q = (a * m) >> 19;
    uint32_t ind = (0u - q) & 7u;
    ind ^= (uint32_t)(ind < 2u);
    bits |= (uint64_t)ind << (3 * i);
  }
  const uint64_t word = (uint64_t)(uint32_t)mx | ((uint64_t)(uint32_t)mn << 8) | (bits << 16);
  return make_uint2((uint32_t)word, (uint32_t)(word >> 32));
}

template <int DXT5>
__global__ void __launch_bounds__(256) encode_kernel(BlockRange r) {
  __shared__ uint8_t s_q[2 * 784];
  load_quant(s_q);
  const uint32_t rel = blockIdx.x * blockDim.x + threadIdx.x;
  if (rel >= r.num_blocks) return;
  const uint32_t b = r.first_block + rel;
  const uint32_t bx = b % r.blocks_wide, by = b / r.blocks_wide;
  const uint32_t *p = r.img + (size_t)(by * 4) * r.width + bx * 4;
  uint32_t blk[16];
#pragma unroll
  for (int y = 0; y < 4; y++) {
    const uint4 row = *reinterpret_cast<const uint4 *>(p + (size_t)y * r.width);
    blk[y * 4 + 0] = row.x;
    blk[y * 4 + 1] = row.y;
    blk[y * 4 + 2] = row.z;
    blk[y * 4 + 3] = row.w;
  }
  const uint2 c = compress_color(blk, s_q);
  if (DXT5) {
    const uint2 a = compress_alpha(blk);
    reinterpret_cast<uint4 *>(r.out)[b] = make_uint4(a.x, a.y, c.x, c.y);
  } else {
    reinterpret_cast<uint2 *>(r.out)[b] = c;
  }
}

// Batch form: persistent workgroups take chunks of the group's blocks from
// a device counter; each block finds its surface through block_surf.
template <int DXT5>
__global__ void batch_kernel(const SurfDesc *surfs, const uint32_t *block_surf, uint32_t total,
                             uint32_t *counter) {
  __shared__ uint32_t s_chunk;
  __shared__ uint8_t s_q[2 * 784];
  load_quant(s_q);
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
    uint32_t blk[16];
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const uint4 row = *reinterpret_cast<const uint4 *>(p + (size_t)y * sd.width);
      blk[y * 4 + 0] = row.x;
      blk[y * 4 + 1] = row.y;
      blk[y * 4 + 2] = row.z;
      blk[y * 4 + 3] = row.w;
    }
    const uint2 c = compress_color(blk, s_q);
    if (DXT5) {
      const uint2 a = compress_alpha(blk);
      reinterpret_cast<uint4 *>(sd.out)[b] = make_uint4(a.x, a.y, c.x, c.y);
    } else {
      reinterpret_cast<uint2 *>(sd.out)[b] = c;
    }
  }
}

}  // namespace dxt

namespace {

int g_dxt_tables[64];
int g_dxt_cus[64];

// stb__InitDXT / stb__PrepareOptTable (stb_dxt.h:121-147, 603-621).
void build_dxt_tables(dxt::Tables &t) {
  auto mul8 = [](int a, int b) {
    const int x = a * b + 128;
    return (x + (x >> 8)) >> 8;
  };
  for (int i = 0; i < 32; i++) t.expand5[i] = (uint8_t)((i << 3) | (i >> 2));
  for (int i = 0; i < 64; i++) t.expand6[i] = (uint8_t)((i << 2) | (i >> 4));
  for (int i = 0; i < 784; i++) {
    const int v = i - 264 < 0 ? 0 : (i - 264 > 255 ? 255 : i - 264);
    t.quant_rb[i] = t.expand5[mul8(v, 31)];
    t.quant_g[i] = t.expand6[mul8(v, 63)];
  }
  for (int pass = 0; pass < 2; pass++) {
    const uint8_t *expand = pass == 0 ? t.expand5 : t.expand6;
    const int size = pass == 0 ? 32 : 64;
    uint8_t(*tab)[2] = pass == 0 ? t.omatch5 : t.omatch6;
    for (int i = 0; i < 256; i++) {
      int best = 256;
      for (int mn = 0; mn < size; mn++)
        for (int mx = 0; mx < size; mx++) {
          const int mine = expand[mn], maxe = expand[mx];
          int err = abs((2 * maxe + mine) / 3 - i);
          err += abs(maxe - mine) * 3 / 100;
          if (err < best) {
            tab[i][0] = (uint8_t)mx;
            tab[i][1] = (uint8_t)mn;
            best = err;
          }
        }
    }
  }
}

}  // namespace

hipError_t dxt_prepare() {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  if (!g_dxt_tables[dev]) {
    dxt::Tables t;
    build_dxt_tables(t);
    e = hipMemcpyToSymbol(HIP_SYMBOL(dxt::kT), &t, sizeof(t));
    if (e != hipSuccess) return e;
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    g_dxt_cus[dev] = cus > 0 ? cus : 1;
    g_dxt_tables[dev] = 1;
  }
  return hipSuccess;
}

hipError_t dxt_encode_batch(const SurfDesc *surfs, const uint32_t *block_surf, uint32_t total,
                            int dxt5, uint32_t *counter, const LaunchTune &tune,
                            hipStream_t stream) {
  if (total == 0) return hipSuccess;
  hipError_t e = dxt_prepare();
  if (e != hipSuccess) return e;
  int dev = 0;
  (void)hipGetDevice(&dev);
  const uint32_t wg = tune.simple_wg;
  uint32_t grid = (uint32_t)g_dxt_cus[dev] * (2048u / wg);
  if (grid > (total + wg - 1) / wg) grid = (total + wg - 1) / wg;
  const int t = timing_begin(stream);
  if (dxt5)
    hipLaunchKernelGGL(dxt::batch_kernel<1>, dim3(grid), dim3(wg), 0, stream, surfs, block_surf,
                       total, counter);
  else
    hipLaunchKernelGGL(dxt::batch_kernel<0>, dim3(grid), dim3(wg), 0, stream, surfs, block_surf,
                       total, counter);
  timing_end(t, kTagDxt, stream);
  return hipGetLastError();
}

hipError_t dxt_encode(const BlockRange &r, int dxt5, hipStream_t stream) {
  if (r.num_blocks == 0) return hipSuccess;
  hipError_t e = dxt_prepare();
  if (e != hipSuccess) return e;
  const uint32_t grid = (r.num_blocks + 255) / 256;
  const int t = timing_begin(stream);
  if (dxt5)
    hipLaunchKernelGGL(dxt::encode_kernel<1>, dim3(grid), dim3(256), 0, stream, r);
  else
    hipLaunchKernelGGL(dxt::encode_kernel<0>, dim3(grid), dim3(256), 0, stream, r);
  timing_end(t, kTagDxt, stream);
  return hipGetLastError();
}

}  // namespace fh
