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
  uint8_t quant_rb[256 + 16];
  uint8_t quant_g[256 + 16];
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

__device__ __forceinline__ int qtab(int ch_, int idx) {
  idx = idx < 0 ? 0 : (idx > 271 ? 271 : idx);
  return ch_ == 1 ? kT.quant_g[idx] : kT.quant_rb[idx];
}

// stb__DitherBlock
__device__ void dither_block(const uint32_t blk[16], uint32_t dst[16]) {
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
      const int d0 = qtab(c, b0 + ((3 * e2[1] + 5 * e2[0]) >> 4) + 8);
      e1[0] = b0 - d0;
      const int d1 = qtab(c, b1 + ((7 * e1[0] + 3 * e2[2] + 5 * e2[1] + e2[0]) >> 4) + 8);
      e1[1] = b1 - d1;
      const int d2 = qtab(c, b2 + ((7 * e1[1] + 3 * e2[3] + 5 * e2[2] + e2[1]) >> 4) + 8);
      e1[2] = b2 - d2;
      const int d3 = qtab(c, b3 + ((7 * e1[2] + 5 * e2[3] + e2[2]) >> 4) + 8);
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

__device__ __forceinline__ int pick_step(int dot, int half, int c0p, int c3p) {
  if (dot < half) return dot < c0p ? 1 : 3;
  return dot < c3p ? 2 : 0;
}

// stb__MatchColorsBlock with dithering.
__device__ uint32_t match_colors(const uint32_t blk[16], const int col[4][3]) {
  const int dirr = col[0][0] - col[1][0];
  const int dirg = col[0][1] - col[1][1];
  const int dirb = col[0][2] - col[1][2];
  int dots[16], stops[4];
#pragma unroll
  for (int i = 0; i < 16; i++) dots[i] = ch(blk[i], 0) * dirr + ch(blk[i], 1) * dirg + ch(blk[i], 2) * dirb;
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
    int st;
    uint32_t lm;
    int dot = (dp[0] << 4) + (3 * e2[1] + 5 * e2[0]);
    st = pick_step(dot, half, c0p, c3p);
    e1[0] = dp[0] - stops[st];
    lm = (uint32_t)st;
    dot = (dp[1] << 4) + (7 * e1[0] + 3 * e2[2] + 5 * e2[1] + e2[0]);
    st = pick_step(dot, half, c0p, c3p);
    e1[1] = dp[1] - stops[st];
    lm |= (uint32_t)st << 2;
    dot = (dp[2] << 4) + (7 * e1[1] + 3 * e2[3] + 5 * e2[2] + e2[1]);
    st = pick_step(dot, half, c0p, c3p);
    e1[2] = dp[2] - stops[st];
    lm |= (uint32_t)st << 4;
    dot = (dp[3] << 4) + (7 * e1[2] + 5 * e2[3] + e2[2]);
    st = pick_step(dot, half, c0p, c3p);
    e1[3] = dp[3] - stops[st];
    lm |= (uint32_t)st << 6;
    mask |= lm << (y * 8);
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
      if (v < minv) minv = v;
      else if (v > maxv) maxv = v;
    }
    mu[c] = (muv + 8) >> 4;
    mn[c] = minv;
    mx[c] = maxv;
  }
  int cov[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int r = ch(blk[i], 0) - mu[0];
    const int g = ch(blk[i], 1) - mu[1];
    const int b = ch(blk[i], 2) - mu[2];
    cov[0] += r * r;
    cov[1] += r * g;
    cov[2] += r * b;
    cov[3] += g * g;
    cov[4] += g * b;
    cov[5] += b * b;
  }
  float covf[6];
#pragma unroll
  for (int i = 0; i < 6; i++) covf[i] = (float)cov[i] / 255.0f;
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
  int mind = 0x7fffffff, maxd = -0x7fffffff;
  uint32_t minp = blk[0], maxp = blk[0];
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int dot = ch(blk[i], 0) * v_r + ch(blk[i], 1) * v_g + ch(blk[i], 2) * v_b;
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
    const int w1tab[4] = {3, 0, 2, 1};
    const int prods[4] = {0x090000, 0x000900, 0x040102, 0x010402};
    int akku = 0, a1r = 0, a1g = 0, a1b = 0, a2r = 0, a2g = 0, a2b = 0;
    uint32_t cm = mask;
#pragma unroll
    for (int i = 0; i < 16; i++, cm >>= 2) {
      const int step = (int)(cm & 3);
      const int w1 = w1tab[step];
      const int r = ch(blk[i], 0), g = ch(blk[i], 1), b = ch(blk[i], 2);
      akku += prods[step];
      a1r += w1 * r;
      a1g += w1 * g;
      a1b += w1 * b;
      a2r += r;
      a2g += g;
      a2b += b;
    }
    a2r = 3 * a2r - a1r;
    a2g = 3 * a2g - a1g;
    a2b = 3 * a2b - a1b;
    const int xx = akku >> 16;
    const int yy = (akku >> 8) & 0xff;
    const int xy = (akku >> 0) & 0xff;
    // 3.0f * 31.0f / 255.0f is evaluated left to right in float
    const float frb = ((3.0f * 31.0f) / 255.0f) / (float)(xx * yy - xy * xy);
    const float fg = (frb * 63.0f) / 31.0f;
    nmax = (uint32_t)sclamp((float)(a1r * yy - a2r * xy) * frb + 0.5f, 0, 31) << 11;
    nmax |= (uint32_t)sclamp((float)(a1g * yy - a2g * xy) * fg + 0.5f, 0, 63) << 5;
    nmax |= (uint32_t)sclamp((float)(a1b * yy - a2b * xy) * frb + 0.5f, 0, 31);
    nmin = (uint32_t)sclamp((float)(a2r * xx - a1r * xy) * frb + 0.5f, 0, 31) << 11;
    nmin |= (uint32_t)sclamp((float)(a2g * xx - a1g * xy) * fg + 0.5f, 0, 63) << 5;
    nmin |= (uint32_t)sclamp((float)(a2b * xx - a1b * xy) * frb + 0.5f, 0, 31);
  }
  min16 = nmin;
  max16 = nmax;
  return old_min != nmin || old_max != nmax;
}

// stb__CompressColorBlock in STB_DXT_DITHER mode (one refinement pass).
__device__ uint2 compress_color(const uint32_t blk[16]) {
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
    dither_block(blk, dblk);
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
    if (a < mn) mn = a;
    else if (a > mx) mx = a;
  }
  const int dist = mx - mn;
  const int dist4 = dist * 4, dist2 = dist * 2;
  int bias = dist < 8 ? dist - 1 : dist / 2 + 2;
  bias -= mn * 7;
  uint64_t bits = 0;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    int a = (int)(blk[i] >> 24) * 7 + bias;
    int t = a >= dist4 ? -1 : 0;
    int ind = t & 4;
    a -= dist4 & t;
    t = a >= dist2 ? -1 : 0;
    ind += t & 2;
    a -= dist2 & t;
    ind += (a >= dist);
    ind = -ind & 7;
    ind ^= (2 > ind);
    bits |= (uint64_t)ind << (3 * i);
  }
  const uint64_t word = (uint64_t)(uint32_t)mx | ((uint64_t)(uint32_t)mn << 8) | (bits << 16);
  return make_uint2((uint32_t)word, (uint32_t)(word >> 32));
}

template <int DXT5>
__global__ void __launch_bounds__(256) encode_kernel(BlockRange r) {
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
  const uint2 c = compress_color(blk);
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
    const uint2 c = compress_color(blk);
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
  for (int i = 0; i < 256 + 16; i++) {
    const int v = i - 8 < 0 ? 0 : (i - 8 > 255 ? 255 : i - 8);
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
