/*
 * BC7 (BPTC) CPU oracle: a plain-C restatement of FasTC's scalar BC7 encoder
 * and decoder.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Arithmetic follows the reference's x86-64 build: IEEE binary32 for the
 * `float` math, binary64 where the reference uses `double`, no FMA
 * contraction (build with -ffp-contract=off), correctly rounded division.
 * Sums are evaluated in the reference's order.
 *
 * Reference files (paths relative to the FasTC tree):
 *   BPTCEncoder/src/Compressor.cpp, CompressionMode.h, RGBAEndpoints.{h,cpp},
 *   BPTCEncoder/src/AnchorTables.h, BCLookupTables.h, Decompressor.cpp,
 *   BPTCEncoder/include/FasTC/Shapes.h,
 *   Base/include/FasTC/{VectorBase.h,MatrixBase.h,MatrixSquare.h,BitStream.h},
 *   Base/src/Image.cpp.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Format tables (BC7 specification; same values as the reference's).       */
/* ------------------------------------------------------------------------ */

/* Two-subset partition masks, bit i set => pixel i in subset 1
 * (Shapes.h:24-33). */
static const uint16_t kShape2[64] = {
  0xcccc, 0x8888, 0xeeee, 0xecc8, 0xc880, 0xfeec, 0xfec8, 0xec80,
  0xc800, 0xffec, 0xfe80, 0xe800, 0xffe8, 0xff00, 0xfff0, 0xf000,
  0xf710, 0x008e, 0x7100, 0x08ce, 0x008c, 0x7310, 0x3100, 0x8cce,
  0x088c, 0x3110, 0x6666, 0x366c, 0x17e8, 0x0ff0, 0x718e, 0x399c,
  0xaaaa, 0xf0f0, 0x5a5a, 0x33cc, 0x3c3c, 0x55aa, 0x9696, 0xa55a,
  0x73ce, 0x13c8, 0x324c, 0x3bdc, 0x6996, 0xc33c, 0x9966, 0x0660,
  0x0272, 0x04e4, 0x4e40, 0x2720, 0xc936, 0x936c, 0x39c6, 0x639c,
  0x9336, 0x9cc6, 0x817e, 0xe718, 0xccf0, 0x0fcc, 0x7744, 0xee22};

/* Three-subset partitions: [0] = pixels not in subset 0, [1] = pixels in
 * subset 2 (Shapes.h:36-53). */
static const uint16_t kShape3[64][2] = {
  {0xfecc, 0xf600}, {0xffc8, 0x7300}, {0xff90, 0x3310}, {0xecce, 0x00ce},
  {0xff00, 0xcc00}, {0xcccc, 0xcc00}, {0xffcc, 0x00cc}, {0xffcc, 0x3300},
  {0xff00, 0xf000}, {0xfff0, 0xf000}, {0xfff0, 0xff00}, {0xcccc, 0x8888},
  {0xeeee, 0x8888}, {0xeeee, 0xcccc}, {0xffec, 0xec80}, {0x739c, 0x7310},
  {0xfec8, 0xc800}, {0x39ce, 0x3100}, {0xfff0, 0xccc0}, {0xfccc, 0x0ccc},
  {0xeeee, 0xee00}, {0xff88, 0x7700}, {0xeec0, 0xcc00}, {0x7730, 0x3300},
  {0x0cee, 0x00cc}, {0xffcc, 0xfc88}, {0x6ff6, 0x0660}, {0xff60, 0x6600},
  {0xcbbc, 0xc88c}, {0xf966, 0xf900}, {0xceec, 0x0cc0}, {0xff10, 0x7310},
  {0xff80, 0xec80}, {0xccce, 0x08ce}, {0xeccc, 0xec80}, {0x6666, 0x4444},
  {0x0ff0, 0x0f00}, {0x6db6, 0x4924}, {0x6bd6, 0x4294}, {0xcf3c, 0x0c30},
  {0xc3fc, 0x03c0}, {0xffaa, 0xff00}, {0xff00, 0x5500}, {0xfcfc, 0xcccc},
  {0xcccc, 0x0c0c}, {0xf6f6, 0x6666}, {0xaffa, 0x0ff0}, {0xfff0, 0x5550},
  {0xfaaa, 0xf000}, {0xeeee, 0x0e0e}, {0xf8f8, 0x8888}, {0xfff0, 0x9990},
  {0xeeee, 0xe00e}, {0x8ff8, 0x8888}, {0xf666, 0xf000}, {0xff00, 0x9900},
  {0xff66, 0xff00}, {0xcccc, 0xc00c}, {0xcffc, 0xcccc}, {0xf000, 0x9000},
  {0x8888, 0x0808}, {0xfefe, 0xeeee}, {0xfffa, 0xfff0}, {0x7bde, 0x7310}};

/* Anchor (fix-up) pixel of subset 1 / subset 2 (AnchorTables.h:22-51). */
static const uint8_t kAnchor2[64] = {
  15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15,
  15, 2,  8,  2,  2,  8,  8,  15, 2,  8,  2,  2,  8,  8,  2,  2,
  15, 15, 6,  8,  2,  8,  15, 15, 2,  8,  2,  2,  2,  15, 15, 6,
  6,  2,  6,  8,  15, 15, 2,  2,  15, 15, 15, 15, 15, 2,  2,  15};
static const uint8_t kAnchor3[2][64] = {
  {3,  3,  15, 15, 8,  3,  15, 15, 8,  8,  6,  6,  6,  5,  3,  3,
   3,  3,  8,  15, 3,  3,  6,  10, 5,  8,  8,  6,  8,  5,  15, 15,
   8,  15, 3,  5,  6,  10, 8,  15, 15, 3,  15, 5,  15, 15, 15, 15,
   3,  15, 5,  5,  5,  8,  5,  10, 5,  10, 8,  13, 15, 12, 3,  3},
  {15, 8,  8,  3,  15, 15, 3,  8,  15, 15, 15, 15, 15, 15, 15, 8,
   15, 8,  15, 3,  15, 8,  15, 8,  3,  15, 6,  10, 15, 15, 10, 8,
   15, 3,  15, 10, 10, 8,  9,  10, 6,  15, 8,  15, 3,  6,  6,  8,
   15, 3,  15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 3,  15, 15, 8}};

/* Interpolation weights for 1..4-bit indices (Compressor.cpp:159-168). */
static const uint32_t kInterp[4][16][2] = {
  {{64, 0}, {33, 31}, {0, 64}},
  {{64, 0}, {43, 21}, {21, 43}, {0, 64}},
  {{64, 0}, {55, 9}, {46, 18}, {37, 27}, {27, 37}, {18, 46}, {9, 55}, {0, 64}},
  {{64, 0}, {60, 4}, {55, 9}, {51, 13}, {47, 17}, {43, 21}, {38, 26}, {34, 30},
   {30, 34}, {26, 38}, {21, 43}, {17, 47}, {13, 51}, {9, 55}, {4, 60}, {0, 64}}};

enum { PB_SHARED = 0, PB_NOTSHARED = 1, PB_NONE = 2 };

/* Per-mode attributes (Compressor.cpp:170-203). */
typedef struct {
  int mode, part_bits, subsets, idx_bits, alpha_idx_bits, color_prec,
      alpha_prec, has_rot, has_idx_mode, pbit_type;
} mode_attr;
static const mode_attr kModes[8] = {
  {0, 4, 3, 3, 0, 4, 0, 0, 0, PB_NOTSHARED},
  {1, 6, 2, 3, 0, 6, 0, 0, 0, PB_SHARED},
  {2, 6, 3, 2, 0, 5, 0, 0, 0, PB_NONE},
  {3, 6, 2, 2, 0, 7, 0, 0, 0, PB_NOTSHARED},
  {4, 0, 1, 2, 3, 5, 6, 1, 1, PB_NONE},
  {5, 0, 1, 2, 2, 7, 8, 1, 0, PB_NONE},
  {6, 0, 1, 4, 0, 7, 7, 0, 0, PB_NOTSHARED},
  {7, 6, 2, 2, 0, 5, 5, 0, 0, PB_NOTSHARED}};

static const int kPBits[4][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}};

/* Solid-block alpha-index watermark (Compressor.cpp:135-138). */
static const uint32_t kWMValues[9] = {0x32b92180, 0x32ba3080, 0x31103200,
                                      0x28103c80, 0x32bb3080, 0x25903600,
                                      0x3530b900, 0x3b32b180, 0x34b5b98};

/* Lookup tables, rebuilt from the generator programs documented in
 * BCLookupTables.h:26-89 (mode 5, 7-bit endpoints, index weight 43/21) and
 * BCLookupTables.h:365-446 (6-bit DXT1-style endpoints). */
static uint8_t gMode5[256][2];
static uint8_t gDxt6[256][2][3];
static int gTablesBuilt = 0;

static void expand_vals(int nbits, uint32_t *vals, int *count) {
  uint32_t last = 0xFFFFFFFFu;
  int n = 0;
  for (uint32_t i = 0; i < 256; i++) {
    uint32_t num = (i >> (8 - nbits)) << (8 - nbits);
    num |= i >> nbits;
    if (num != last) {
      last = num;
      vals[n++] = num;
    }
  }
  *count = n;
}

static void build_tables(void) {
  if (gTablesBuilt) return;
  uint32_t v[256];
  int nv;
  expand_vals(7, v, &nv);
  for (uint32_t i = 0; i < 256; i++) {
    uint32_t md = 0xFFFFFFFFu, mj = 0, mk = 0;
    for (int j = 0; j < nv; j++)
      for (int k = 0; k < nv; k++) {
        uint32_t c = (43 * v[j] + 21 * v[k] + 32) >> 6;
        uint32_t d = i > c ? i - c : c - i;
        if (d < md) { md = d; mj = j; mk = k; }
      }
    gMode5[i][0] = (uint8_t)(v[mj] >> 1);
    gMode5[i][1] = (uint8_t)(v[mk] >> 1);
  }
  expand_vals(6, v, &nv);
  for (uint32_t i = 0; i < 256; i++) {
    uint32_t e[2][4] = {{1, 0, 0, 0xFFFFFFFFu}, {0, 0, 0, 0xFFFFFFFFu}};
    for (int w = 0; w < 2; w++) {
      uint32_t md = 0xFFFFFFFFu, mj = 0, mk = 0;
      for (int j = 0; j < nv; j++)
        for (int k = j; k < nv; k++) {
          uint32_t c = w == 0 ? (v[j] + v[k]) / 2 : (2 * v[j] + v[k]) / 3;
          uint32_t d = i > c ? i - c : c - i;
          if (d < md) { md = d; mj = j; mk = k; }
        }
      e[w][1] = v[mj];
      e[w][2] = v[mk];
      e[w][3] = md;
    }
    int first = e[1][3] > e[0][3] ? 0 : 1;
    for (int t = 0; t < 2; t++) {
      int w = t == 0 ? first : 1 - first;
      gDxt6[i][t][0] = (uint8_t)e[w][0];
      gDxt6[i][t][1] = (uint8_t)(e[w][1] >> 2);
      gDxt6[i][t][2] = (uint8_t)(e[w][2] >> 2);
    }
  }
  gTablesBuilt = 1;
}

const uint8_t *ora_table_mode5(void) { build_tables(); return &gMode5[0][0]; }
const uint8_t *ora_table_dxt6(void) { build_tables(); return &gDxt6[0][0][0]; }

void ora_table_partitions(uint16_t shape2[64], uint16_t shape3[128], uint8_t anchor2[64],
                          uint8_t anchor3[128], uint32_t interp[128]) {
  memcpy(shape2, kShape2, sizeof(kShape2));
  memcpy(shape3, kShape3, sizeof(kShape3));
  memcpy(anchor2, kAnchor2, sizeof(kAnchor2));
  memcpy(anchor3, kAnchor3, sizeof(kAnchor3));
  memcpy(interp, kInterp, sizeof(kInterp));
}

/* ------------------------------------------------------------------------ */
/* Vector helpers: FasTC::VectorBase<float,4> semantics (VectorBase.h).     */
/* ------------------------------------------------------------------------ */

typedef struct { float v[4]; } v4;

static v4 v4_set(float a, float b, float c, float d) {
  v4 r; r.v[0] = a; r.v[1] = b; r.v[2] = c; r.v[3] = d; return r;
}
static v4 v4_splat(float a) { return v4_set(a, a, a, a); }
/* RGBAVector(idx, pixel) (RGBAEndpoints.h:64-72). */
static v4 v4_pixel(uint32_t p) {
  return v4_set((float)(p & 0xFF), (float)((p >> 8) & 0xFF),
                (float)((p >> 16) & 0xFF), (float)((p >> 24) & 0xFF));
}
static v4 v4_sub(v4 a, v4 b) { for (int i = 0; i < 4; i++) a.v[i] -= b.v[i]; return a; }
static v4 v4_add(v4 a, v4 b) { for (int i = 0; i < 4; i++) a.v[i] += b.v[i]; return a; }
/* VectorBase::Dot (VectorBase.h:96-101): sum from 0 in channel order. */
static float v4_dot(v4 a, v4 b) {
  float s = 0;
  for (int i = 0; i < 4; i++) s += a.v[i] * b.v[i];
  return s;
}
/* ScalarMultiply / ScalarDivide (VectorBase.h:186-192, 289-294). */
static v4 v4_mul(v4 a, float s) { for (int i = 0; i < 4; i++) a.v[i] = a.v[i] * s; return a; }
static v4 v4_div(v4 a, float s) { for (int i = 0; i < 4; i++) a.v[i] = a.v[i] / s; return a; }
static int v4_eq(v4 a, v4 b) {
  return a.v[0] == b.v[0] && a.v[1] == b.v[1] && a.v[2] == b.v[2] && a.v[3] == b.v[3];
}
/* VectorBase::Length: float(sqrt(long double(LengthSq))) (VectorBase.h:104-105). */
static float v4_len(v4 a) {
  return (float)sqrtl((long double)v4_dot(a, a));
}
static v4 v4_normalize(v4 a) {
  float len = v4_len(a);
  for (int i = 0; i < 4; i++) a.v[i] /= len;
  return a;
}

float ora_length(const float v[4]) {
  v4 a = v4_set(v[0], v[1], v[2], v[3]);
  return v4_len(a);
}

/* ------------------------------------------------------------------------ */
/* Quantisation (RGBAEndpoints.cpp:126-177).                                */
/* ------------------------------------------------------------------------ */

static uint32_t count_bits(uint8_t n) {
  uint32_t c = 0;
  while (n) { n &= (uint8_t)(n - 1); c++; }
  return c;
}

static uint8_t sad8(uint8_t a, uint8_t b) { return a > b ? (uint8_t)(a - b) : (uint8_t)(b - a); }

uint8_t ora_quantize_channel(uint8_t val, uint8_t mask, int pbit) {
  if (mask == 0xFF) return val;
  if (mask == 0x00) return 0xFF;
  uint32_t prec = count_bits(mask);
  const uint32_t step = 1u << (8 - prec);
  uint32_t lval = val & mask;
  uint32_t hval = lval + step;
  if (pbit >= 0) {
    prec++;
    lval |= (uint32_t)(!!pbit) << (8 - prec);
    hval |= (uint32_t)(!!pbit) << (8 - prec);
  }
  if (lval > val) {
    lval -= step;
    hval -= step;
  }
  lval |= lval >> prec;
  hval |= hval >> prec;
  if (sad8(val, (uint8_t)lval) < sad8(val, (uint8_t)hval))
    return (uint8_t)lval;
  return (uint8_t)hval;
}

/* uint32(c + 0.5) with the add in double (RGBAEndpoints.cpp:169-172). */
static uint32_t round_half_up(float c) { return (uint32_t)((double)c + 0.5); }

static uint32_t to_pixel(v4 p, uint32_t mask, int pbit) {
  uint32_t r = ora_quantize_channel((uint8_t)(round_half_up(p.v[0]) & 0xFF), mask & 0xFF, pbit);
  uint32_t g = ora_quantize_channel((uint8_t)(round_half_up(p.v[1]) & 0xFF), (mask >> 8) & 0xFF, pbit);
  uint32_t b = ora_quantize_channel((uint8_t)(round_half_up(p.v[2]) & 0xFF), (mask >> 16) & 0xFF, pbit);
  uint32_t a = ora_quantize_channel((uint8_t)(round_half_up(p.v[3]) & 0xFF), (mask >> 24) & 0xFF, pbit);
  return r | (g << 8) | (b << 16) | (a << 24);
}

uint32_t ora_to_pixel(const float v[4], uint32_t mask, int pbit) {
  return to_pixel(v4_set(v[0], v[1], v[2], v[3]), mask, pbit);
}

/* ------------------------------------------------------------------------ */
/* RGBACluster (RGBAEndpoints.h:104-227).                                   */
/* ------------------------------------------------------------------------ */

typedef struct {
  uint32_t n, nparts, part, shape;
  v4 avg, mn, mx;
  v4 pts[16];
  uint32_t pix[16];
  uint8_t map[16];
} clus;

/* GetSubsetForIndex (Shapes.h:55-79). */
static int subset_for(int idx, int shape, int nsubs) {
  if (nsubs == 2) return !!((1 << idx) & kShape2[shape]);
  if (nsubs == 3) {
    if ((1 << idx) & kShape3[shape][0]) return 1 + !!((1 << idx) & kShape3[shape][1]);
    return 0;
  }
  return 0;
}

/* GetAnchorIndexForSubset (AnchorTables.h:55-83). */
static int anchor_for(int subset, int shape, int nsubs) {
  if (subset == 1) return nsubs == 2 ? kAnchor2[shape] : kAnchor3[0][shape];
  if (subset == 2) return kAnchor3[1][shape];
  return 0;
}

static void clus_init(clus *c, const uint32_t px[16]) {
  c->n = 0;
  c->avg = v4_splat(0.0f);
  c->mn = v4_splat(FLT_MAX);
  c->mx = v4_splat(-FLT_MAX);
  c->nparts = 1; c->part = 0; c->shape = 0;
  for (uint32_t i = 0; i < 16; i++) {
    v4 p = v4_pixel(px[i]);
    c->avg = v4_add(c->avg, p);
    c->map[c->n] = (uint8_t)i;
    c->pix[c->n] = to_pixel(p, 0xFFFFFFFFu, -1);
    c->pts[c->n++] = p;
    for (int k = 0; k < 4; k++) {
      if (p.v[k] < c->mn.v[k]) c->mn.v[k] = p.v[k];
      if (p.v[k] > c->mx.v[k]) c->mx.v[k] = p.v[k];
    }
  }
  c->avg = v4_div(c->avg, (float)c->n);
}

/* Recalculate (RGBAEndpoints.h:205-226). */
static void clus_set_partition(clus *c, uint32_t part) {
  c->part = part;
  c->n = 0;
  c->avg = v4_splat(0.0f);
  c->mn = v4_splat(FLT_MAX);
  c->mx = v4_splat(-FLT_MAX);
  uint32_t map = 0;
  for (uint32_t idx = 0; idx < 16; idx++) {
    if ((uint32_t)subset_for((int)idx, (int)c->shape, (int)c->nparts) != part) continue;
    c->n++;
    c->avg = v4_add(c->avg, c->pts[idx]);
    c->map[map++] = (uint8_t)idx;
    for (int k = 0; k < 4; k++) {
      if (c->pts[idx].v[k] < c->mn.v[k]) c->mn.v[k] = c->pts[idx].v[k];
      if (c->pts[idx].v[k] > c->mx.v[k]) c->mx.v[k] = c->pts[idx].v[k];
    }
  }
  c->avg = v4_div(c->avg, (float)c->n);
}

static v4 clus_pt(const clus *c, uint32_t i) { return c->pts[c->map[i]]; }
static uint32_t clus_pix(const clus *c, uint32_t i) { return c->pix[c->map[i]]; }

/* RGBACluster::QuantizedError<nBuckets> (RGBAEndpoints.cpp:190-310). */
static double qerr(const clus *c, v4 p1, v4 p2, int nb, uint32_t mask, v4 metric,
                   const int *pbits, uint8_t *indices) {
  const int prec = nb == 4 ? 2 : (nb == 8 ? 3 : 4);
  const uint32_t(*iv)[2] = kInterp[prec - 1];
  uint32_t qp1, qp2;
  if (pbits) {
    qp1 = to_pixel(p1, mask, pbits[0]);
    qp2 = to_pixel(p2, mask, pbits[1]);
  } else {
    qp1 = to_pixel(p1, mask, -1);
    qp2 = to_pixel(p2, mask, -1);
  }
  const v4 uqp1 = v4_pixel(qp1), uqp2 = v4_pixel(qp2);
  const float uqplsq = v4_dot(v4_sub(uqp1, uqp2), v4_sub(uqp1, uqp2));
  const v4 uqpdir = v4_sub(uqp2, uqp1);
  float total = 0.0f;
  if (uqplsq == 0) {
    for (uint32_t i = 0; i < c->n; i++) {
      const uint32_t pixel = clus_pix(c, i);
      v4 ev;
      for (int k = 0; k < 4; k++) {
        uint32_t a = (qp1 >> (8 * k)) & 0xFF, b = (qp2 >> (8 * k)) & 0xFF;
        uint32_t ip = ((a * iv[0][0]) + (b * iv[0][1]) + 32) >> 6 & 0xFF;
        uint8_t d = sad8((uint8_t)((pixel >> (8 * k)) & 0xFF), (uint8_t)ip);
        ev.v[k] = (float)d * metric.v[k];
      }
      total += v4_dot(ev, ev);
      if (indices) indices[i] = 0;
    }
    return total;
  }
  for (uint32_t i = 0; i < c->n; i++) {
    const v4 pt = clus_pt(c, i);
    const float pct = v4_dot(v4_sub(pt, uqp1), uqpdir) / uqplsq;
    int32_t j1 = (int32_t)floorf(pct * (float)(nb - 1));
    int32_t j2 = (int32_t)ceilf(pct * (float)(nb - 1));
    if (j1 < 0) j1 = 0;
    if (j1 > nb - 1) j1 = nb - 1;
    if (j2 > nb - 1) j2 = nb - 1;
    const uint32_t pixel = clus_pix(c, i);
    float minErr = FLT_MAX;
    uint8_t best = 0;
    int32_t j = j1;
    do {
      v4 ev;
      for (int k = 0; k < 4; k++) {
        uint32_t a = (qp1 >> (8 * k)) & 0xFF, b = (qp2 >> (8 * k)) & 0xFF;
        uint32_t ip = ((a * iv[j][0]) + (b * iv[j][1]) + 32) >> 6 & 0xFF;
        uint8_t d = sad8((uint8_t)((pixel >> (8 * k)) & 0xFF), (uint8_t)ip);
        ev.v[k] = (float)d * metric.v[k];
      }
      float e = v4_dot(ev, ev);
      if (e < minErr) {
        minErr = e;
        best = (uint8_t)j;
      } else if (e > minErr) {
        break;
      }
    } while (++j <= j2);
    total += minErr;
    if (indices) indices[i] = best;
  }
  return total;
}

double ora_quantized_error_w(const uint32_t px[16], const uint8_t *sel, int n,
                             const float p1[4], const float p2[4], int nbuckets,
                             uint32_t mask, const int *pbits, const float w[4],
                             uint8_t *indices) {
  clus c;
  clus_init(&c, px);
  if (sel) {
    c.n = (uint32_t)n;
    for (int i = 0; i < n; i++) c.map[i] = sel[i];
  }
  return qerr(&c, v4_set(p1[0], p1[1], p1[2], p1[3]), v4_set(p2[0], p2[1], p2[2], p2[3]),
              nbuckets, mask, v4_set(w[0], w[1], w[2], w[3]), pbits, indices);
}

double ora_quantized_error(const uint32_t px[16], const uint8_t *sel, int n,
                           const float p1[4], const float p2[4], int nbuckets,
                           uint32_t mask, const int *pbits, uint8_t *indices) {
  static const float ones[4] = {1.0f, 1.0f, 1.0f, 1.0f};
  return ora_quantized_error_w(px, sel, n, p1, p2, nbuckets, mask, pbits, ones, indices);
}

/* MatrixSquare<float,4>::PowerMethod (MatrixSquare.h:44-105). */
static int power_method(const float m[16], v4 *eig, float *eigval) {
  const int kMax = 5;
  int iters = 0;
  v4 b;
  const float norm = 1.0f / sqrtf(4.0f);
  for (int i = 0; i < 4; i++) b.v[i] = norm;
  int bad = 0, fixed = 0;
  while (!fixed && ++iters < kMax) {
    v4 nb;
    for (int j = 0; j < 4; j++) {
      nb.v[j] = 0;
      for (int i = 0; i < 4; i++) nb.v[j] += m[j * 4 + i] * b.v[i];
    }
    const float len = v4_len(nb);
    if ((double)len < 1e-10) {
      if (bad) {
        *eig = b;
        if (eigval) *eigval = 0.0f;
        return iters;
      }
      for (int i = 0; i < 2; i++) b.v[i] = 1;
      b = v4_normalize(b);
      bad = 1;
      continue;
    }
    nb = v4_normalize(nb);
    if (fabs((double)(1.0f - v4_dot(b, nb))) < 1e-8) fixed = 1;
    b = nb;
  }
  *eig = b;
  if (eigval) {
    v4 r;
    for (int j = 0; j < 4; j++) {
      r.v[j] = 0;
      for (int i = 0; i < 4; i++) r.v[j] += m[j * 4 + i] * b.v[i];
    }
    *eigval = v4_len(r) / v4_len(b);
  }
  return iters;
}

int ora_power_method(const float m[16], float eig[4], float *eigval) {
  v4 e;
  int it = power_method(m, &e, eigval);
  for (int i = 0; i < 4; i++) eig[i] = e.v[i];
  return it;
}

/* RGBACluster::GetPrincipalAxis with eigOne = eigTwo = NULL
 * (RGBAEndpoints.cpp:327-428).  Reproduces the read of default-valued
 * (-1,-1,-1,-1) entries past the unique-point count in the collinearity
 * test (RGBAEndpoints.cpp:364-371).  Pinned by the float64 eigenvector of
 * the covariance and by known answers of the collinear shortcut with and
 * without that padding (tests/test_oracle_fit.py). */
static void principal_axis(const clus *c, v4 *axis) {
  v4 to[16];
  for (uint32_t i = 0; i < c->n; i++) to[i] = v4_sub(clus_pt(c, i), c->avg);
  v4 upts[16];
  for (int i = 0; i < 16; i++) upts[i] = v4_splat(-1.0f);
  uint32_t nu = 0;
  for (uint32_t i = 0; i < c->n; i++) {
    int has = 0;
    for (uint32_t j = 0; j < nu; j++)
      if (v4_eq(upts[j], clus_pt(c, i))) has = 1;
    if (!has) upts[nu++] = clus_pt(c, i);
  }
  if (nu == 1) {
    *axis = v4_splat(0.0f);
    return;
  }
  {
    v4 dir = v4_normalize(v4_sub(upts[1], upts[0]));
    int collinear = 1;
    for (uint32_t i = 2; i < c->n; i++) {
      v4 v = v4_sub(upts[i], upts[0]);
      if (fabs(fabs((double)v4_dot(v, dir)) - (double)v4_len(v)) > 1e-7) {
        collinear = 0;
        break;
      }
    }
    if (collinear) {
      *axis = dir;
      return;
    }
  }
  float cov[16];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j <= i; j++) {
      float sum = 0.0f;
      for (uint32_t k = 0; k < c->n; k++) sum += to[k].v[i] * to[k].v[j];
      cov[i * 4 + j] = sum / 3.0f;
      cov[j * 4 + i] = cov[i * 4 + j];
    }
  power_method(cov, axis, NULL);
}

/* ------------------------------------------------------------------------ */
/* CompressionMode (CompressionMode.h:63-307, Compressor.cpp:212-1371).     */
/* ------------------------------------------------------------------------ */

typedef struct {
  const mode_attr *a;
  int rot, idx_mode;
  uint32_t sa_steps;
  v4 metric;
  /* annealing RNG */
  int rng_kind;
  uint32_t *lcg;
  uint32_t key, ctr;
  uint64_t seed, block_id;
  ora_fit_trace *trace; /* optional: intermediate values of the RGB fit */
} cmode;

typedef struct {
  v4 p1[3], p2[3];
  uint8_t idx[3][16];
  uint8_t aidx[16];
  uint8_t pbc[3];
  int rot, idx_mode;
  int shape;
} params;

static int m_opaque(const cmode *m) { return m->a->mode < 4; }
static int m_rot(const cmode *m) { return m->a->has_rot ? m->rot : 0; }
static int m_bits_idx(const cmode *m, int im) {
  if (im < 0) im = m->idx_mode;
  return im == 0 ? m->a->idx_bits : m->a->alpha_idx_bits;
}
static int m_bits_alpha(const cmode *m, int im) {
  if (im < 0) im = m->idx_mode;
  return im == 0 ? m->a->alpha_idx_bits : m->a->idx_bits;
}

/* GetQuantizationMask (CompressionMode.h:212-232). */
static uint32_t m_qmask(const cmode *m) {
  const int32_t seed = (int32_t)0x80000000u;
  const uint32_t cbits = (uint32_t)m->a->color_prec - 1;
  const uint32_t abits = (uint32_t)m->a->alpha_prec - 1;
  if (m->a->alpha_prec > 0) {
    return ((uint32_t)(seed >> (24 + cbits)) & 0xFF) | ((uint32_t)(seed >> (16 + cbits)) & 0xFF00) |
           ((uint32_t)(seed >> (8 + cbits)) & 0xFF0000) | ((uint32_t)(seed >> abits) & 0xFF000000u);
  }
  return (((uint32_t)(seed >> (24 + cbits)) & 0xFF) | ((uint32_t)(seed >> (16 + cbits)) & 0xFF00) |
          ((uint32_t)(seed >> (8 + cbits)) & 0xFF0000)) & 0x00FFFFFF;
}

static int m_npbc(const cmode *m) {
  return m->a->pbit_type == PB_SHARED ? 2 : (m->a->pbit_type == PB_NOTSHARED ? 4 : 1);
}
static const int *m_pbc(const cmode *m, int idx) {
  if (m->a->pbit_type == PB_SHARED) return idx ? kPBits[3] : kPBits[0];
  if (m->a->pbit_type == PB_NOTSHARED) return kPBits[idx % 4];
  return kPBits[0];
}

/* kErrorMetrics (Compressor.cpp:205-208): per-channel error weights of the
 * uniform (0) and nonuniform (1) metrics.  QuantizedError multiplies each
 * channel's absolute difference by its weight before squaring. */
static v4 metric_weights(uint32_t metric) {
  if (metric == 1) return v4_set(sqrtf(0.3f), sqrtf(0.56f), sqrtf(0.11f), 1.0f);
  return v4_splat(1.0f);
}

/* GetErrorMetric (CompressionMode.h:196-205): the weights with the rotated
 * channel's weight swapped with alpha's; m->metric holds them unrotated. */
static v4 m_metric(const cmode *m) {
  const v4 w = m->metric;
  switch (m_rot(m)) {
    case 1: return v4_set(w.v[3], w.v[1], w.v[2], w.v[0]);
    case 2: return v4_set(w.v[0], w.v[3], w.v[2], w.v[1]);
    case 3: return v4_set(w.v[0], w.v[1], w.v[3], w.v[2]);
    default: return w;
  }
}

/* Counter-based annealing stream (shared definition with the HIP kernels). */
static uint32_t mix32(uint32_t x) {
  x ^= x >> 16; x *= 0x7feb352du;
  x ^= x >> 15; x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
uint32_t ora_counter_rng(uint32_t key, uint32_t n) { return mix32(key + n * 0x9E3779B9u); }
uint32_t ora_fit_key(uint64_t seed, uint64_t block_id, uint32_t fit_tag) {
  uint32_t k = mix32((uint32_t)seed ^ 0x2545F491u);
  k = mix32(k ^ (uint32_t)(seed >> 32));
  k = mix32(k ^ (uint32_t)block_id);
  k = mix32(k ^ (uint32_t)(block_id >> 32));
  return mix32(k ^ fit_tag);
}

/* fastrand (Compressor.cpp:358-362).  With glibc RAND_MAX = 2^31-1 the mask
 * keeps all 16 bits of (g_seed >> 16). */
static uint32_t m_fastrand(cmode *m) {
  if (m->rng_kind == ORA_RNG_LCG) {
    *m->lcg = 214013u * *m->lcg + 2531011u;
    return (*m->lcg >> 16) & 0x7FFFFFFFu;
  }
  return ora_counter_rng(m->key, m->ctr++) >> 16;
}

/* frand (Compressor.cpp:506-519). */
static float m_frand(cmode *m) {
  const uint16_t r = (uint16_t)m_fastrand(m);
  const uint32_t bits = (127u << 23) | (((uint32_t)r << 8) | ((uint32_t)r >> 7));
  float f;
  memcpy(&f, &bits, 4);
  return f - 1.0f;
}

/* ClampEndpoints (RGBAEndpoints.cpp:436-441). */
static float clampf_ref(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }
static void clamp_endpoints(v4 *p1, v4 *p2) {
  for (int i = 0; i < 4; i++) {
    p1->v[i] = clampf_ref(p1->v[i], 0.0f, 255.0f);
    p2->v[i] = clampf_ref(p2->v[i], 0.0f, 255.0f);
  }
}

/* ClampEndpointsToGrid (Compressor.cpp:212-250).  Pinned by a float64 brute
 * force over the p-bit combinations (tests/test_oracle_fit.py). */
static void clamp_to_grid(const cmode *m, v4 *p1, v4 *p2, uint8_t *best_pbc) {
  const int npbc = m_npbc(m);
  const int has_pb = npbc > 1;
  const uint32_t qmask = m_qmask(m);
  clamp_endpoints(p1, p2);
  float min_dist = FLT_MAX;
  v4 bp1 = v4_splat(-1.0f), bp2 = v4_splat(-1.0f);
  for (int i = 0; i < npbc; i++) {
    uint32_t q1, q2;
    if (has_pb) {
      q1 = to_pixel(*p1, qmask, m_pbc(m, i)[0]);
      q2 = to_pixel(*p2, qmask, m_pbc(m, i)[1]);
    } else {
      q1 = to_pixel(*p1, qmask, -1);
      q2 = to_pixel(*p2, qmask, -1);
    }
    v4 n1 = v4_pixel(q1), n2 = v4_pixel(q2);
    v4 d1 = v4_sub(n1, *p1), d2 = v4_sub(n2, *p2);
    float dist = v4_dot(d1, d1) + v4_dot(d2, d2);
    if (dist < min_dist) {
      min_dist = dist;
      bp1 = n1;
      bp2 = n2;
      *best_pbc = (uint8_t)i;
    }
  }
  *p1 = bp1;
  *p2 = bp2;
}

/* CompressSingleColor (Compressor.cpp:252-353). */
static double single_color(const cmode *m, v4 p, v4 *p1, v4 *p2, uint8_t *best_pbc) {
  const uint32_t pixel = to_pixel(p, 0xFFFFFFFFu, -1);
  float best = FLT_MAX;
  for (int pbi = 0; pbi < m_npbc(m); pbi++) {
    const int *combo = m_pbc(m, pbi);
    uint32_t dist[4] = {0, 0, 0, 0};
    uint32_t bi[4], bj[4];
    memset(bi, 0xFF, sizeof(bi));
    memset(bj, 0xFF, sizeof(bj));
    for (uint32_t ci = 0; ci < 4; ci++) {
      const uint8_t val = (uint8_t)((pixel >> (ci * 8)) & 0xFF);
      int nbits = m->a->color_prec;
      if (ci == 3) nbits = m->a->alpha_prec;
      if (nbits == 0) {
        bi[ci] = bj[ci] = 0xFF;
        uint32_t d = (uint32_t)(0xFF - val);
        dist[ci] = dist[ci] > d ? dist[ci] : d;
        continue;
      }
      const int nposs = 1 << nbits;
      int vh_tab[256], vl_tab[256];
      const int havepbit = m->a->pbit_type != PB_NONE;
      if (havepbit) nbits++;
      for (int i = 0; i < nposs; i++) {
        int vh = i, vl = i;
        if (havepbit) {
          vh <<= 1; vl <<= 1;
          vh |= combo[1];
          vl |= combo[0];
        }
        vh_tab[i] = vh << (8 - nbits);
        vh_tab[i] |= vh_tab[i] >> nbits;
        vl_tab[i] = vl << (8 - nbits);
        vl_tab[i] |= vl_tab[i] >> nbits;
      }
      const int bpi = m_bits_idx(m, -1) - 1;
      const uint32_t w0 = kInterp[bpi][1][0], w1 = kInterp[bpi][1][1];
      uint32_t bestd = 0xFF;
      for (int i = 0; bestd > 0 && i < nposs; i++)
        for (int j = 0; bestd > 0 && j < nposs; j++) {
          const uint32_t v1 = (uint32_t)vl_tab[i], v2 = (uint32_t)vh_tab[j];
          const uint32_t combo_v = (w0 * v1 + (w1 * v2) + 32) >> 6;
          const uint32_t err = combo_v > val ? combo_v - val : val - combo_v;
          if (err < bestd) {
            bestd = err;
            bi[ci] = v1;
            bj[ci] = v2;
          }
        }
      dist[ci] = bestd > dist[ci] ? bestd : dist[ci];
    }
    float error = 0.0f;
    for (uint32_t i = 0; i < 4; i++) {
      float e = (float)dist[i] * m->metric.v[i];
      error += e * e;
    }
    if (error < best) {
      best = error;
      *best_pbc = (uint8_t)pbi;
      for (uint32_t ci = 0; ci < 4; ci++) {
        p1->v[ci] = (float)bi[ci];
        p2->v[ci] = (float)bj[ci];
      }
    }
  }
  return best;
}

double ora_single_color_m(int mode, int idx_mode, uint32_t metric, const float p[4],
                          float p1[4], float p2[4], int *pbc) {
  cmode m;
  memset(&m, 0, sizeof(m));
  m.a = &kModes[mode & 7];
  m.idx_mode = idx_mode;
  m.metric = metric_weights(metric);
  v4 pv, a, b;
  for (int k = 0; k < 4; k++) pv.v[k] = p[k];
  uint8_t bp = 0;
  const double e = single_color(&m, pv, &a, &b, &bp);
  for (int k = 0; k < 4; k++) {
    p1[k] = a.v[k];
    p2[k] = b.v[k];
  }
  *pbc = bp;
  return e;
}

double ora_single_color(int mode, int idx_mode, const float p[4], float p1[4], float p2[4],
                        int *pbc) {
  return ora_single_color_m(mode, idx_mode, 0, p, p1, p2, pbc);
}

/* ChangePointForDir{Without,With}PbitChange (Compressor.cpp:364-418). */
static void change_point(v4 *v, uint32_t dir, const float step[4]) {
  for (int k = 0; k < 4; k++) {
    if ((dir >> k) & 1) v->v[k] -= step[k];
    else v->v[k] += step[k];
  }
}
static void change_point_pbit(v4 *v, uint32_t dir, uint32_t oldp, const float step[4]) {
  for (int k = 0; k < 4; k++) {
    int bit = (dir >> k) & 1;
    if (bit && oldp == 0) v->v[k] -= step[k];
    else if (!bit && oldp == 1) v->v[k] += step[k];
  }
}

/* PickBestNeighboringEndpoints (Compressor.cpp:426-498) with the single
 * visited state the caller maintains. */
static void pick_neighbor(cmode *m, v4 p1, v4 p2, int cur_pbc, v4 *np1, v4 *np2,
                          int *npbc, const v4 *vp1, const v4 *vp2, const int *vpbc,
                          int nvisited) {
  float step[4] = {1.0f * (float)(1 << (8 - m->a->color_prec)),
                   1.0f * (float)(1 << (8 - m->a->color_prec)),
                   1.0f * (float)(1 << (8 - m->a->color_prec)),
                   1.0f * (float)(1 << (8 - m->a->alpha_prec))};
  if (m_opaque(m)) step[(m_rot(m) + 3) % 4] = 0.0f;
  const int has_pb = m->a->pbit_type != PB_NONE;
  if (has_pb) {
    if (m->a->pbit_type == PB_SHARED) *npbc = (cur_pbc + 1) % 2;
    else *npbc = 3 - cur_pbc;
  }
  int visited = 1;
  int guard = -1;
  while (visited && ++guard < 16) {
    for (int pt = 0; pt < 2; pt++) {
      const v4 p = pt ? p1 : p2;
      v4 *np = pt ? np1 : np2;
      *np = p;
      if (has_pb) {
        const uint32_t rdir = m_fastrand(m) % 16;
        const uint32_t pbit = (uint32_t)m_pbc(m, cur_pbc)[pt];
        change_point_pbit(np, rdir, pbit, step);
      } else {
        change_point(np, m_fastrand(m) % 16, step);
      }
      for (int i = 0; i < 4; i++) {
        float x = np->v[i];
        x = x < 0.0f ? 0.0f : x;      /* std::max(np[i], 0.0f) */
        x = 255.0f < x ? 255.0f : x;  /* std::min(.., 255.0f) */
        np->v[i] = x;
      }
    }
    visited = 0;
    for (int i = 0; i < nvisited; i++)
      visited = visited || (v4_eq(vp1[i], *np1) && v4_eq(vp2[i], *np2) && vpbc[i] == *npbc);
  }
}

/* AcceptNewEndpointError (Compressor.cpp:524-536). */
static int accept_new(cmode *m, double new_err, double old_err, float temp) {
  if (new_err < old_err) return 1;
  const double p = exp(((double)0.1f * (old_err - new_err)) / (double)temp);
  const double r = (double)m_frand(m);
  return r < p;
}

/* QuantizedError point evaluations of the annealing loops run by this
 * thread (one per point of the subset per step). */
static _Thread_local uint64_t g_anneal_points;

uint64_t ora_anneal_point_count(int reset) {
  const uint64_t n = g_anneal_points;
  if (reset) g_anneal_points = 0;
  return n;
}

/* OptimizeEndpointsForCluster (Compressor.cpp:538-630). */
static double optimize_endpoints(cmode *m, const clus *c, v4 *p1, v4 *p2, uint8_t *best_idx,
                                 uint8_t *best_pbc) {
  const int nb = 1 << m_bits_idx(m, -1);
  const uint32_t qmask = m_qmask(m);
  double cur = qerr(c, *p1, *p2, nb, qmask, m_metric(m), m_pbc(m, *best_pbc), best_idx);
  int cur_pbc = *best_pbc;
  double best = cur;
  uint32_t q1, q2;
  if (m->a->pbit_type != PB_NONE) {
    q1 = to_pixel(*p1, qmask, m_pbc(m, *best_pbc)[0]);
    q2 = to_pixel(*p2, qmask, m_pbc(m, *best_pbc)[1]);
  } else {
    q1 = to_pixel(*p1, qmask, -1);
    q2 = to_pixel(*p2, qmask, -1);
  }
  *p1 = v4_pixel(q1);
  *p2 = v4_pixel(q2);
  v4 bp1 = *p1, bp2 = *p2;
  v4 vp1 = *p1, vp2 = *p2;
  int vpbc = cur_pbc;
  const int max_energy = (int)m->sa_steps;
  for (int energy = 0; best > 0 && energy < max_energy; energy++) {
    float temp = (float)energy / (float)(max_energy - 1);
    uint8_t idx[16];
    v4 np1, np2;
    int npbc = 0;
    pick_neighbor(m, *p1, *p2, cur_pbc, &np1, &np2, &npbc, &vp1, &vp2, &vpbc, 1);
    double err = qerr(c, np1, np2, nb, qmask, m_metric(m), m_pbc(m, npbc), idx);
    g_anneal_points += c->n;
    if (accept_new(m, err, cur, temp)) {
      cur = err;
      *p1 = np1;
      *p2 = np2;
      cur_pbc = npbc;
    }
    if (err < best) {
      memcpy(best_idx, idx, 16);
      bp1 = np1;
      bp2 = np2;
      *best_pbc = (uint8_t)npbc;
      best = err;
      vp1 = np1;
      vp2 = np2;
      vpbc = npbc;
      energy = 0;
    }
  }
  *p1 = bp1;
  *p2 = bp2;
  return best;
}

static void trace_v4(float out[4], v4 v) {
  for (int k = 0; k < 4; k++) out[k] = v.v[k];
}

/* CompressCluster, RGB variant (Compressor.cpp:921-1094).  Pinned at each
 * stage through m->trace (tests/test_oracle_fit.py): the axis, the k-means
 * fixed point, the least-squares solve, the grid clamp, and exact fits of
 * two-colour subsets. */
static double compress_cluster_rgb(cmode *m, const clus *c, v4 *p1, v4 *p2, uint8_t *best_idx,
                                   uint8_t *best_pbc) {
  ora_fit_trace *tr = m->trace;
  if (tr) memset(tr, 0, sizeof(*tr));
  if (v4_eq(c->mx, c->mn)) {
    double e = single_color(m, clus_pt(c, 0), p1, p2, best_pbc);
    for (uint32_t i = 0; i < c->n; i++) best_idx[i] = 1;
    if (tr) {
      tr->path = ORA_FIT_SAME_POINT;
      trace_v4(tr->single, clus_pt(c, 0));
    }
    return (double)c->n * e;
  }
  const uint32_t nb = 1u << m_bits_idx(m, -1);
  v4 axis;
  principal_axis(c, &axis);
  float mindp = FLT_MAX, maxdp = -FLT_MAX;
  for (uint32_t i = 0; i < c->n; i++) {
    float dp = v4_dot(v4_sub(clus_pt(c, i), c->avg), axis);
    if (dp < mindp) mindp = dp;
    if (dp > maxdp) maxdp = dp;
  }
  *p1 = v4_add(c->avg, v4_mul(axis, mindp));
  *p2 = v4_add(c->avg, v4_mul(axis, maxdp));
  clamp_endpoints(p1, p2);
  if (tr) {
    trace_v4(tr->axis, axis);
    trace_v4(tr->start1, *p1);
    trace_v4(tr->start2, *p2);
    tr->nbuckets = nb;
  }

  v4 pts[16];
  uint32_t npts[16];
  for (uint32_t i = 0; i < nb; i++) {
    float s = (float)i / (float)(nb - 1);
    pts[i] = v4_add(v4_mul(*p1, 1.0f - s), v4_mul(*p2, s));
  }
  uint32_t bidx[16] = {0};
  int fixed = 0;
  while (!fixed) {
    v4 newp[16];
    for (uint32_t i = 0; i < c->n; i++) {
      int minb = -1;
      float mind = FLT_MAX;
      for (uint32_t j = 0; j < nb; j++) {
        v4 v = v4_sub(clus_pt(c, i), pts[j]);
        float d = v4_dot(v, v);
        if (d < mind) {
          mind = d;
          minb = (int)j;
        }
      }
      bidx[i] = (uint32_t)minb;
    }
    for (uint32_t i = 0; i < nb; i++) {
      npts[i] = 0;
      newp[i] = v4_splat(0.0f);
      for (uint32_t j = 0; j < c->n; j++)
        if (bidx[j] == i) {
          npts[i]++;
          newp[i] = v4_add(newp[i], clus_pt(c, j));
        }
      if (npts[i] != 0) newp[i] = v4_div(newp[i], (float)npts[i]);
    }
    fixed = 1;
    for (uint32_t i = 0; i < nb; i++)
      if (!v4_eq(pts[i], newp[i])) fixed = 0;
    for (uint32_t i = 0; i < nb; i++) pts[i] = newp[i];
    if (tr) tr->iterations++;
  }
  if (tr) {
    for (uint32_t i = 0; i < nb; i++) {
      trace_v4(tr->centroid[i], pts[i]);
      tr->count[i] = npts[i];
    }
    for (uint32_t i = 0; i < c->n; i++) tr->bucket[i] = (uint8_t)bidx[i];
  }
  int nfilled = 0, last = -1;
  for (uint32_t i = 0; i < nb; i++)
    if (npts[i] > 0) {
      nfilled++;
      last = (int)i;
    }
  if (nfilled == 1) {
    double e = single_color(m, pts[last], p1, p2, best_pbc);
    for (uint32_t i = 0; i < c->n; i++) best_idx[i] = 1;
    if (tr) {
      tr->path = ORA_FIT_ONE_BUCKET;
      trace_v4(tr->single, pts[last]);
    }
    return (double)c->n * e;
  }
  float asq = 0.0f, bsq = 0.0f, ab = 0.0f;
  v4 ax = v4_splat(0.0f), bx = v4_splat(0.0f);
  for (uint32_t i = 0; i < nb; i++) {
    const v4 x = pts[i];
    const int n = (int)npts[i];
    const float fbi = (float)(nb - 1 - i);
    const float fb = (float)(nb - 1);
    const float fi = (float)i;
    const float fn = (float)n;
    const float a = fbi / fb;
    const float b = fi / fb;
    asq += fn * a * a;
    bsq += fn * b * b;
    ab += fn * a * b;
    ax = v4_add(ax, v4_mul(v4_mul(x, a), fn));
    bx = v4_add(bx, v4_mul(v4_mul(x, b), fn));
  }
  const float f = 1.0f / (asq * bsq - ab * ab);
  *p1 = v4_mul(v4_sub(v4_mul(ax, bsq), v4_mul(bx, ab)), f);
  *p2 = v4_mul(v4_sub(v4_mul(bx, asq), v4_mul(ax, ab)), f);
  if (tr) {
    tr->path = ORA_FIT_LEAST_SQUARES;
    trace_v4(tr->lsq1, *p1);
    trace_v4(tr->lsq2, *p2);
  }
  clamp_to_grid(m, p1, p2, best_pbc);
  if (tr) {
    trace_v4(tr->grid1, *p1);
    trace_v4(tr->grid2, *p2);
    tr->pbit_combo = *best_pbc;
  }
  return optimize_endpoints(m, c, p1, p2, best_idx, best_pbc);
}

/* std::min(255.0f, std::max(0.0f, x)) (Compressor.cpp:871-872). */
static float clamp_alpha(float x) {
  float t = 0.0f < x ? x : 0.0f;
  return t < 255.0f ? t : 255.0f;
}

/* CompressCluster, alpha/rotation variant (Compressor.cpp:632-919).  Pinned
 * by known answers of the stale-statistics quirk, of the equal-alpha mode-4
 * table branch and of the alpha k-means (tests/test_oracle_fit.py). */
static double compress_cluster_alpha(cmode *m, const clus *c, v4 *p1, v4 *p2, uint8_t *best_idx,
                                     uint8_t *alpha_idx) {
  if (v4_eq(c->mx, c->mn)) {
    uint8_t dummy = 0;
    double e = single_color(m, clus_pt(c, 0), p1, p2, &dummy);
    for (uint32_t i = 0; i < c->n; i++) {
      best_idx[i] = 1;
      alpha_idx[i] = 1;
    }
    return (double)c->n * e;
  }
  /* The copy keeps the source cluster's avg/min/max; only the points are
   * rotated and given A = 255 (Compressor.cpp:664-695). */
  clus rc = *c;
  float avals[16] = {0};
  float amin = FLT_MAX, amax = -FLT_MAX;
  for (uint32_t i = 0; i < rc.n; i++) {
    v4 *v = &rc.pts[rc.map[i]];
    int r = m_rot(m);
    if (r >= 1 && r <= 3) {
      float t = v->v[r - 1];
      v->v[r - 1] = v->v[3];
      v->v[3] = t;
    }
    avals[i] = v->v[3];
    v->v[3] = 255.0f;
    amin = avals[i] < amin ? avals[i] : amin;
    amax = amax < avals[i] ? avals[i] : amax;
  }
  uint8_t dummy_pbc = 0;
  v4 rp1, rp2;
  double rgb_err = compress_cluster_rgb(m, &rc, &rp1, &rp2, best_idx, &dummy_pbc);
  ora_fit_trace *tr = m->trace;
  if (tr) tr->rgb_error = rgb_err;

  float a1 = amin, a2 = amax;
  double alpha_err = DBL_MAX;
  const int abits = m_bits_alpha(m, -1);
  const uint32_t(*iv)[2] = kInterp[abits - 1];
  const float weight = m_metric(m).v[3];
  const uint32_t nb = 1u << abits;
  if (a1 == a2) {
    const uint8_t a1be = (uint8_t)a1, a2be = (uint8_t)a2;
    if (m->a->mode == 5) {
      for (int i = 0; i < 16; i++) alpha_idx[i] = 0;
      alpha_err = 0.0;
    } else {
      build_tables();
      if (gDxt6[a1be][0][0]) {
        a1 = (float)((gDxt6[a1be][1][1] << 2) | (gDxt6[a1be][0][1] >> 4));
        a2 = (float)((gDxt6[a2be][1][2] << 2) | (gDxt6[a2be][0][1] >> 4));
      } else {
        a1 = (float)((gDxt6[a1be][0][1] << 2) | (gDxt6[a1be][0][1] >> 4));
        a2 = (float)((gDxt6[a2be][0][2] << 2) | (gDxt6[a2be][0][1] >> 4));
      }
      uint8_t fill = m->idx_mode == 1 ? 1 : 2;
      for (int i = 0; i < 16; i++) alpha_idx[i] = fill;
      const uint32_t i0 = iv[alpha_idx[0]][0], i1 = iv[alpha_idx[0]][1];
      const uint32_t a1i = (uint32_t)a1, a2i = (uint32_t)a2;
      const uint8_t ip = (uint8_t)((((a1i * i0) + (a2i * i1) + 32) >> 6) & 0xFF);
      float pe = weight * (float)(a1be > ip ? a1be - ip : ip - a1be);
      pe *= pe;
      alpha_err = 16 * (double)pe;
    }
  } else {
    float vals[8];
    memset(vals, 0, sizeof(vals));
    uint32_t buckets[16];
    memset(buckets, 0, sizeof(buckets));
    for (uint32_t i = 0; i < nb; i++) {
      const float fi = (float)i;
      const float fb = (float)(nb - 1);
      vals[i] = amin + (fi / fb) * (amax - amin);
    }
    for (uint32_t i = 0; i < 16; i++) {
      float md = 255.0f;
      for (uint32_t j = 0; j < nb; j++) {
        float d = fabsf(avals[i] - vals[j]);
        if (d < md) {
          md = d;
          buckets[i] = j;
        }
      }
    }
    float npts[8];
    int fixed = 0;
    while (!fixed) {
      memset(npts, 0, sizeof(npts));
      float avg[8];
      memset(avg, 0, sizeof(avg));
      for (uint32_t i = 0; i < nb; i++) {
        for (uint32_t j = 0; j < 16; j++)
          if (buckets[j] == i) {
            avg[i] += avals[j];
            npts[i] += 1.0f;
          }
        if (npts[i] > 0.0f) avg[i] /= npts[i];
      }
      fixed = 1;
      for (uint32_t i = 0; i < nb; i++) fixed = fixed && (avg[i] == vals[i]);
      memcpy(vals, avg, sizeof(vals));
      for (uint32_t i = 0; i < 16; i++) {
        float md = 255.0f;
        for (uint32_t j = 0; j < nb; j++) {
          float d = fabsf(avals[i] - vals[j]);
          if (d < md) {
            md = d;
            buckets[i] = j;
          }
        }
      }
    }
    float asq = 0.0f, bsq = 0.0f, ab = 0.0f, ax = 0.0f, bx = 0.0f;
    for (uint32_t i = 0; i < nb; i++) {
      const float fbi = (float)(nb - 1 - i);
      const float fb = (float)(nb - 1);
      const float fi = (float)i;
      float a = fbi / fb;
      float b = fi / fb;
      float n = npts[i];
      float x = vals[i];
      asq += n * a * a;
      bsq += n * b * b;
      ab += n * a * b;
      ax += x * a * n;
      bx += x * b * n;
    }
    float f = 1.0f / (asq * bsq - ab * ab);
    a1 = f * (ax * bsq - bx * ab);
    a2 = f * (bx * asq - ax * ab);
    a1 = clamp_alpha(a1);
    a2 = clamp_alpha(a2);
    const int8_t mask_seed = -0x7F;
    const uint8_t amask = (uint8_t)(mask_seed >> (m->a->alpha_prec - 1));
    const uint8_t a1b = ora_quantize_channel((uint8_t)a1, amask, -1);
    const uint8_t a2b = ora_quantize_channel((uint8_t)a2, amask, -1);
    if (tr) {
      tr->alpha_path = ORA_FIT_LEAST_SQUARES;
      for (uint32_t i = 0; i < nb; i++) {
        tr->alpha_centroid[i] = vals[i];
        tr->alpha_count[i] = (uint32_t)npts[i];
      }
      for (uint32_t i = 0; i < 16; i++) tr->alpha_bucket[i] = (uint8_t)buckets[i];
      tr->alpha_lsq[0] = a1;
      tr->alpha_lsq[1] = a2;
      tr->alpha_q[0] = a1b;
      tr->alpha_q[1] = a2b;
    }
    alpha_err = 0.0;
    for (uint32_t i = 0; i < 16; i++) {
      uint8_t val = (uint8_t)avals[i];
      float min_e = FLT_MAX;
      int bb = -1;
      for (uint32_t j = 0; j < nb; j++) {
        const uint32_t a1i = a1b, a2i = a2b;
        const uint8_t ip = (uint8_t)((((a1i * iv[j][0]) + (a2i * iv[j][1]) + 32) >> 6) & 0xFF);
        float pe = weight * (float)(val > ip ? val - ip : ip - val);
        pe *= pe;
        if (pe < min_e) {
          min_e = pe;
          bb = (int)j;
        }
      }
      alpha_err += min_e;
      alpha_idx[i] = (uint8_t)bb;
    }
  }
  for (int i = 0; i < 4; i++) {
    p1->v[i] = i == 3 ? a1 : rp1.v[i];
    p2->v[i] = i == 3 ? a2 : rp2.v[i];
  }
  if (tr) tr->alpha_error = alpha_err;
  return rgb_err + alpha_err;
}

/* Test hooks into the cluster fits.  A cluster is the subset `subset` of
 * partition shape `shape` of the mode's subset count (SetShapeIndex +
 * SetPartition, RGBAEndpoints.h:205-226). */
static void fit_setup(cmode *m, clus *c, const uint32_t px[16], int mode, int shape, int subset,
                      int rot, int idx_mode, uint32_t metric) {
  memset(m, 0, sizeof(*m));
  m->a = &kModes[mode & 7];
  m->metric = metric_weights(metric);
  m->rot = rot;
  m->idx_mode = idx_mode;
  clus_init(c, px);
  c->shape = (uint32_t)(m->a->subsets > 1 ? shape : 0);
  c->nparts = (uint32_t)m->a->subsets;
  clus_set_partition(c, (uint32_t)subset);
}

void ora_principal_axis(const uint32_t px[16], int nparts, int shape, int subset, float axis[4]) {
  clus c;
  clus_init(&c, px);
  c.shape = (uint32_t)(nparts > 1 ? shape : 0);
  c.nparts = (uint32_t)nparts;
  clus_set_partition(&c, (uint32_t)subset);
  v4 a;
  principal_axis(&c, &a);
  trace_v4(axis, a);
}

void ora_clamp_to_grid(int mode, float p1[4], float p2[4], int *pbc) {
  cmode m;
  memset(&m, 0, sizeof(m));
  m.a = &kModes[mode & 7];
  v4 a = v4_set(p1[0], p1[1], p1[2], p1[3]), b = v4_set(p2[0], p2[1], p2[2], p2[3]);
  uint8_t bp = 0xFF;
  clamp_to_grid(&m, &a, &b, &bp);
  trace_v4(p1, a);
  trace_v4(p2, b);
  *pbc = bp;
}

double ora_fit_rgb_m(const uint32_t px[16], int mode, int shape, int subset, int idx_mode,
                     uint32_t metric, ora_fit_trace *tr, float p1[4], float p2[4],
                     uint8_t idx[16], int *pbc) {
  cmode m;
  clus c;
  fit_setup(&m, &c, px, mode, shape, subset, 0, idx_mode, metric);
  m.trace = tr;
  v4 a, b;
  uint8_t bp = 0;
  memset(idx, 0, 16);
  const double e = compress_cluster_rgb(&m, &c, &a, &b, idx, &bp);
  trace_v4(p1, a);
  trace_v4(p2, b);
  *pbc = bp;
  return e;
}

double ora_fit_alpha_m(const uint32_t px[16], int mode, int rot, int idx_mode, uint32_t metric,
                       ora_fit_trace *tr, float p1[4], float p2[4], uint8_t idx[16],
                       uint8_t aidx[16]) {
  cmode m;
  clus c;
  fit_setup(&m, &c, px, mode, 0, 0, rot, idx_mode, metric);
  m.trace = tr;
  v4 a, b;
  memset(idx, 0, 16);
  memset(aidx, 0, 16);
  const double e = compress_cluster_alpha(&m, &c, &a, &b, idx, aidx);
  trace_v4(p1, a);
  trace_v4(p2, b);
  return e;
}

double ora_fit_rgb(const uint32_t px[16], int mode, int shape, int subset, int idx_mode,
                   ora_fit_trace *tr, float p1[4], float p2[4], uint8_t idx[16], int *pbc) {
  return ora_fit_rgb_m(px, mode, shape, subset, idx_mode, 0, tr, p1, p2, idx, pbc);
}

double ora_fit_alpha(const uint32_t px[16], int mode, int rot, int idx_mode, ora_fit_trace *tr,
                     float p1[4], float p2[4], uint8_t idx[16], uint8_t aidx[16]) {
  return ora_fit_alpha_m(px, mode, rot, idx_mode, 0, tr, p1, p2, idx, aidx);
}

/* Fit tag for the counter RNG: identifies (mode, shape entry, subset,
 * rotation, index mode) within a block. */
static uint32_t fit_tag(int mode, int shape_entry, int subset, int rot, int im) {
  /* shape entries 0..4: the two low bits, then bit 10 */
  return (uint32_t)mode | ((uint32_t)(shape_entry & 3) << 3) | ((uint32_t)subset << 5) |
         ((uint32_t)rot << 7) | ((uint32_t)im << 9) | ((uint32_t)(shape_entry >> 2) << 10);
}

static void m_rekey(cmode *m, int shape_entry, int subset) {
  if (m->rng_kind != ORA_RNG_COUNTER) return;
  m->key = ora_fit_key(m->seed, m->block_id,
                       fit_tag(m->a->mode, shape_entry, subset, m->rot, m->idx_mode));
  m->ctr = 0;
}

/* CompressionMode::Compress (Compressor.cpp:1300-1371). */
static double mode_compress(cmode *m, params *pr, int shape_idx, int shape_entry, clus *c) {
  const int nsub = m->a->subsets;
  memset(pr, 0, sizeof(*pr));
  memset(pr->idx, 0xFF, sizeof(pr->idx));
  memset(pr->aidx, 0xFF, sizeof(pr->aidx));
  memset(pr->pbc, 0xFF, sizeof(pr->pbc));
  pr->rot = -1;
  pr->idx_mode = -1;
  pr->shape = shape_idx;
  for (int i = 0; i < 3; i++) pr->p1[i] = pr->p2[i] = v4_splat(-1.0f);
  double total = 0.0;
  for (int cidx = 0; cidx < nsub; cidx++) {
    uint8_t idx[16] = {0};
    clus_set_partition(c, (uint32_t)cidx);
    if (m->a->has_rot) {
      uint8_t aidx[16];
      double best = DBL_MAX;
      for (int rot = 0; rot < 4; rot++) {
        m->rot = rot;
        const int nim = m->a->mode == 4 ? 2 : 1;
        for (int im = 0; im < nim; im++) {
          m->idx_mode = im;
          m_rekey(m, shape_entry, cidx);
          v4 v1, v2;
          double e = compress_cluster_alpha(m, c, &v1, &v2, idx, aidx);
          if (e < best) {
            best = e;
            memcpy(pr->idx[cidx], idx, 16);
            memcpy(pr->aidx, aidx, 16);
            pr->rot = rot;
            pr->idx_mode = im;
            pr->p1[cidx] = v1;
            pr->p2[cidx] = v2;
          }
        }
      }
      total += best;
    } else {
      m_rekey(m, shape_entry, cidx);
      total += compress_cluster_rgb(m, c, &pr->p1[cidx], &pr->p2[cidx], idx, &pr->pbc[cidx]);
      int k = 0;
      for (int i = 0; i < 16; i++)
        if (subset_for(i, shape_idx, nsub) == cidx) pr->idx[cidx][i] = idx[k++];
    }
  }
  return total;
}

/* BitStream::WriteBits (BitStream.h:65-91): LSB-first into 128 bits. */
typedef struct {
  uint8_t *p;
  int nbits;
} bitw;
static void bw_write(bitw *w, uint32_t val, int n) {
  for (int i = 0; i < n; i++) {
    if (w->nbits >= 128) return;
    const int byte = w->nbits >> 3, bit = w->nbits & 7;
    uint32_t b = i < 32 ? (val >> i) & 1 : 0;
    w->p[byte] = (uint8_t)((w->p[byte] & ~(1u << bit)) | (b << bit));
    w->nbits++;
  }
}

/* CompressionMode::Pack (Compressor.cpp:1096-1298).  Pinned for every mode by
 * an independent BC7-specification reader in the tests (fields, anchors,
 * rotation) and by the internal error of fits equalling their decoded
 * error.  The p-bits are written from the combo as chosen, also when the
 * anchor fix-up swapped the two endpoints (Compressor.cpp:1149-1170 vs
 * 1234-1241). */
static void mode_pack(const cmode *m, params *pr, uint8_t out[16]) {
  bitw w = {out, 0};
  const int mode = m->a->mode;
  const int nsub = m->a->subsets;
  bw_write(&w, 1u << mode, mode + 1);
  bw_write(&w, (uint32_t)pr->shape, m->a->part_bits);
  bw_write(&w, (uint32_t)pr->rot, m->a->has_rot ? 2 : 0);
  bw_write(&w, (uint32_t)pr->idx_mode, m->a->has_idx_mode ? 1 : 0);
  const uint32_t qmask = m_qmask(m);
  uint32_t px1[3], px2[3];
  for (int i = 0; i < nsub; i++) {
    if (m->a->pbit_type == PB_NONE) {
      px1[i] = to_pixel(pr->p1[i], qmask, -1);
      px2[i] = to_pixel(pr->p2[i], qmask, -1);
    } else {
      px1[i] = to_pixel(pr->p1[i], qmask, m_pbc(m, pr->pbc[i])[0]);
      px2[i] = to_pixel(pr->p2[i], qmask, m_pbc(m, pr->pbc[i])[1]);
    }
  }
  for (int s = 0; s < nsub; s++) {
    const int anchor = anchor_for(s, pr->shape, nsub);
    const int nab = m_bits_alpha(m, pr->idx_mode);
    const int nib = m_bits_idx(m, pr->idx_mode);
    if (pr->idx[s][anchor] >> (nib - 1)) {
      uint32_t t = px1[s]; px1[s] = px2[s]; px2[s] = t;
      const int nv = 1 << nib;
      for (int i = 0; i < 16; i++) pr->idx[s][i] = (uint8_t)((nv - 1) - pr->idx[s][i]);
      if (m->a->has_rot) {
        const int nav = 1 << nab;
        for (int i = 0; i < 16; i++) pr->aidx[i] = (uint8_t)((nav - 1) - pr->aidx[i]);
      }
    }
    if (m->a->has_rot && (pr->aidx[anchor] >> (nab - 1)) > 0) {
      uint32_t a1 = px1[s] >> 24, a2 = px2[s] >> 24;
      px1[s] = (px1[s] & 0x00FFFFFFu) | (a2 << 24);
      px2[s] = (px2[s] & 0x00FFFFFFu) | (a1 << 24);
      const int nav = 1 << nab;
      for (int i = 0; i < 16; i++) pr->aidx[i] = (uint8_t)((nav - 1) - pr->aidx[i]);
    }
  }
  const int cp = m->a->color_prec, ap = m->a->alpha_prec;
  for (int ch = 0; ch < 3; ch++)
    for (int i = 0; i < nsub; i++) {
      bw_write(&w, ((px1[i] >> (8 * ch)) & 0xFF) >> (8 - cp), cp);
      bw_write(&w, ((px2[i] >> (8 * ch)) & 0xFF) >> (8 - cp), cp);
    }
  for (int i = 0; i < nsub; i++) {
    bw_write(&w, ap ? ((px1[i] >> 24) & 0xFF) >> (8 - ap) : 0, ap);
    bw_write(&w, ap ? ((px2[i] >> 24) & 0xFF) >> (8 - ap) : 0, ap);
  }
  if (m->a->pbit_type != PB_NONE) {
    for (int s = 0; s < nsub; s++) {
      const int *pb = m_pbc(m, pr->pbc[s]);
      bw_write(&w, (uint32_t)pb[0], 1);
      if (m->a->pbit_type != PB_SHARED) bw_write(&w, (uint32_t)pb[1], 1);
    }
  }
  if (m->a->has_idx_mode && pr->idx_mode == 1) {
    for (int i = 0; i < 16; i++) bw_write(&w, pr->aidx[i], i == 0 ? 1 : 2);
    for (int i = 0; i < 16; i++) bw_write(&w, pr->idx[0][i], i == 0 ? 2 : 3);
  } else {
    const int nbi = m_bits_idx(m, pr->idx_mode);
    for (int i = 0; i < 16; i++) {
      const int s = subset_for(i, pr->shape, nsub);
      const int anchor = anchor_for(s, pr->shape, nsub);
      bw_write(&w, pr->idx[s][i], i == anchor ? nbi - 1 : nbi);
    }
    if (m->a->has_rot) {
      const int nba = m_bits_alpha(m, pr->idx_mode);
      for (int i = 0; i < 16; i++) bw_write(&w, pr->aidx[i], i == 0 ? nba - 1 : nba);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Block level (Compressor.cpp:1405-1458, 1626-1860).                       */
/* ------------------------------------------------------------------------ */

static const uint32_t kTwoPartModes = (1u << 1) | (1u << 3) | (1u << 7);
static const uint32_t kThreePartModes = (1u << 0) | (1u << 2);
static const uint32_t kAlphaModes = (1u << 4) | (1u << 5) | (1u << 6) | (1u << 7);

/* EstimateTwo/ThreeClusterError (Compressor.cpp:1626-1655). */
static double estimate(const clus *c, int nb, v4 w) {
  v4 v = v4_sub(c->mx, c->mn);
  if (v4_dot(v, v) == 0) return 0.0;
  double e = 0.0001;
  e += qerr(c, c->mn, c->mx, nb, 0xFFFFFFFFu, w, NULL, NULL);
  return e;
}

typedef struct {
  uint32_t nshapes;
  int sh_idx[5], sh_parts[5];  /* the entries CompressClusters tries (at most 5) */
  uint32_t modes;
} shape_sel;

/* BoxSelection (Compressor.cpp:1670-1750); its estimates weigh the channels
 * by the unrotated metric (:1634, 1650). */
static shape_sel box_selection(const uint32_t px[16], uint32_t metric) {
  const v4 w = metric_weights(metric);
  shape_sel r;
  r.modes = 0xFF;
  memset(r.sh_idx, 0, sizeof(r.sh_idx));
  memset(r.sh_parts, 0, sizeof(r.sh_parts));
  int opaque = 1;
  for (int i = 0; i < 16; i++) opaque = opaque && (((px[i] >> 24) & 0xFF) >= 250);
  double best[2] = {DBL_MAX, DBL_MAX};
  clus c;
  clus_init(&c, px);
  r.nshapes = 1;
  for (int i = 0; i < 64; i++) {
    c.shape = (uint32_t)i; c.nparts = 2;
    double err = 0.0;
    for (int ci = 0; ci < 2; ci++) {
      clus_set_partition(&c, (uint32_t)ci);
      err += estimate(&c, 8, w);
    }
    if (err < best[0]) {
      best[0] = err;
      r.sh_idx[0] = i;
      r.sh_parts[0] = 2;
    }
    if (err < 1e-9) {
      r.modes = kTwoPartModes;
      return r;
    }
  }
  if (!opaque) {
    r.modes &= kAlphaModes;
    return r;
  }
  r.modes &= ~((1u << 4) | (1u << 5));
  r.nshapes++;
  for (int i = 0; i < 64; i++) {
    c.shape = (uint32_t)i; c.nparts = 3;
    double err = 0.0;
    for (int ci = 0; ci < 3; ci++) {
      clus_set_partition(&c, (uint32_t)ci);
      err += estimate(&c, 4, w);
    }
    if (err < best[1]) {
      best[1] = err;
      r.sh_idx[1] = i;
      r.sh_parts[1] = 3;
    }
    if (err < 1e-9) {
      r.modes = kThreePartModes;
      return r;
    }
  }
  return r;
}

void ora_bc7_box_selection_m(const uint32_t px[16], uint32_t metric, uint32_t *nshapes,
                             uint32_t *shape2, uint32_t *shape3, uint32_t *modes) {
  shape_sel s = box_selection(px, metric);
  *nshapes = s.nshapes;
  *shape2 = (uint32_t)s.sh_idx[0];
  *shape3 = (uint32_t)s.sh_idx[1];
  *modes = s.modes;
}

void ora_bc7_box_selection(const uint32_t px[16], uint32_t *nshapes, uint32_t *shape2,
                           uint32_t *shape3, uint32_t *modes) {
  ora_bc7_box_selection_m(px, 0, nshapes, shape2, shape3, modes);
}

/* CompressOptimalColorBC7 (Compressor.cpp:1424-1458). */
static void write_solid(uint32_t pixel, uint8_t out[16], uint32_t wm_index) {
  build_tables();
  bitw w = {out, 0};
  bw_write(&w, 1u << 5, 6);
  bw_write(&w, 0, 2);
  for (int ch = 0; ch < 3; ch++) {
    const uint8_t v = (uint8_t)((pixel >> (8 * ch)) & 0xFF);
    bw_write(&w, gMode5[v][0], 7);
    bw_write(&w, gMode5[v][1], 7);
  }
  const uint8_t a = (uint8_t)(pixel >> 24);
  bw_write(&w, a, 8);
  bw_write(&w, a, 8);
  bw_write(&w, 0xaaaaaaabu, 31);
  bw_write(&w, kWMValues[wm_index % 9], 31);
}

/* WriteTransparentBlock (Compressor.cpp:1416-1421). */
static void write_transparent(uint8_t out[16]) {
  bitw w = {out, 0};
  bw_write(&w, 1u << 6, 7);
  bw_write(&w, 0, 121);
}

static int all_one_color(const uint32_t px[16]) {
  for (int i = 1; i < 16; i++)
    if (px[i] != px[0]) return 0;
  return 1;
}

/* CompressClusters (Compressor.cpp:1752-1817). */
static void compress_clusters(const shape_sel *sel, const uint32_t px[16],
                              const ora_bc7_settings *s, uint64_t block_id, uint8_t out[16],
                              ora_bc7_block_stats *st) {
  clus c;
  clus_init(&c, px);
  double best_err = DBL_MAX;
  static const int order[8] = {0, 2, 1, 3, 7, 4, 5, 6};
  int best_mode = 8;
  params best_pr;
  memset(&best_pr, 0, sizeof(best_pr));
  uint32_t modes = sel->modes;
  uint32_t nsh = sel->nshapes < 5 ? sel->nshapes : 5;
  if (nsh == 0) {
    nsh = 1;
    modes &= ~(kTwoPartModes | kThreePartModes);
  }
  for (int mi = 0; mi < 8; mi++) {
    const int mode = order[mi];
    if ((modes & (1u << mode)) == 0) continue;
    for (uint32_t si = 0; si < nsh; si++) {
      const uint32_t nparts = (uint32_t)kModes[mode].subsets;
      if (nparts != 1 && nparts != (uint32_t)sel->sh_parts[si]) continue;
      if (sel->sh_idx[si] >= 16 && mode == 0) continue;
      const int idx = sel->sh_idx[si];
      c.shape = (uint32_t)idx;
      c.nparts = nparts;
      cmode m;
      memset(&m, 0, sizeof(m));
      m.a = &kModes[mode];
      m.sa_steps = s->quality;
      m.metric = metric_weights(s->error_metric);
      m.rng_kind = s->rng_kind;
      m.lcg = s->lcg_state;
      m.seed = s->seed;
      m.block_id = block_id;
      params pr;
      double e = mode_compress(&m, &pr, idx, (int)si, &c);
      if (st) {
        if (st->mode_error[mode] < 0 || e < st->mode_error[mode]) st->mode_error[mode] = e;
      }
      if (e < best_err) {
        best_err = e;
        best_mode = mode;
        best_pr = pr;
      }
    }
  }
  memset(out, 0, 16);
  /* No mode tried (the block-mode mask removed every candidate): the
   * reference asserts (Compressor.cpp:1811) and is undefined without
   * asserts; an all-zero block (reserved mode, decodes to zero) is written. */
  if (best_mode < 8) {
    cmode pm;
    memset(&pm, 0, sizeof(pm));
    pm.a = &kModes[best_mode];
    pm.metric = v4_splat(1.0f);
    mode_pack(&pm, &best_pr, out);
  }
  if (st) {
    st->mode = best_mode;
    st->best_error = best_err;
  }
}

static void decode_block(const uint8_t blk[16], uint32_t px[16], int spec);

/* One candidate, packed: CompressionMode::Compress and Pack for a fixed
 * (mode, slot) as in ora_bc7_exhaustive_label, with `quality` annealing
 * steps on the counter stream keyed by (seed, block_id, mode, shape entry,
 * subset, rotation, index mode).  Returns the internal error. */
static double encode_candidate(const uint32_t px[16], int mode, int slot, int entry,
                               uint32_t quality, uint64_t seed, uint64_t block_id,
                               uint8_t out[16]) {
  clus c;
  clus_init(&c, px);
  cmode m;
  memset(&m, 0, sizeof(m));
  m.a = &kModes[mode];
  m.metric = v4_splat(1.0f);
  m.sa_steps = quality;
  m.rng_kind = ORA_RNG_COUNTER;
  m.seed = seed;
  m.block_id = block_id;
  params pr;
  double e;
  if (m.a->has_rot) {
    memset(&pr, 0, sizeof(pr));
    memset(pr.idx, 0xFF, sizeof(pr.idx));
    memset(pr.pbc, 0xFF, sizeof(pr.pbc));
    clus_set_partition(&c, 0);
    m.rot = mode == 4 ? slot >> 1 : slot;
    m.idx_mode = mode == 4 ? slot & 1 : 0;
    m_rekey(&m, entry, 0);
    e = compress_cluster_alpha(&m, &c, &pr.p1[0], &pr.p2[0], pr.idx[0], pr.aidx);
    pr.rot = m.rot;
    pr.idx_mode = m.idx_mode;
    pr.shape = 0;
  } else {
    c.shape = (uint32_t)(m.a->subsets > 1 ? slot : 0);
    c.nparts = (uint32_t)m.a->subsets;
    e = mode_compress(&m, &pr, (int)c.shape, entry, &c);
  }
  memset(out, 0, 16);
  cmode pm;
  memset(&pm, 0, sizeof(pm));
  pm.a = &kModes[mode];
  pm.metric = v4_splat(1.0f);
  mode_pack(&pm, &pr, out);
  return e;
}

void ora_bc7_pack(int mode, int shape, int rot, int idx_mode, const float ep[3][2][4],
                  const uint8_t pbc[3], const uint8_t idx[16], const uint8_t aidx[16],
                  uint8_t out[16]) {
  cmode m;
  memset(&m, 0, sizeof(m));
  m.a = &kModes[mode & 7];
  m.metric = v4_splat(1.0f);
  params pr;
  memset(&pr, 0, sizeof(pr));
  const int nsub = m.a->subsets;
  pr.shape = nsub > 1 ? shape : 0;
  pr.rot = m.a->has_rot ? rot : 0;
  pr.idx_mode = m.a->has_idx_mode ? idx_mode : 0;
  for (int s = 0; s < nsub; s++) {
    pr.p1[s] = v4_set(ep[s][0][0], ep[s][0][1], ep[s][0][2], ep[s][0][3]);
    pr.p2[s] = v4_set(ep[s][1][0], ep[s][1][1], ep[s][1][2], ep[s][1][3]);
    pr.pbc[s] = pbc[s];
  }
  for (int i = 0; i < 16; i++) {
    pr.idx[subset_for(i, pr.shape, nsub)][i] = idx[i];
    pr.aidx[i] = aidx[i];
  }
  memset(out, 0, 16);
  mode_pack(&m, &pr, out);
}

double ora_bc7_encode_candidate(const uint32_t px[16], int mode, int slot, uint8_t out[16]) {
  return encode_candidate(px, mode, slot, 0, 0, 0, 0, out);
}

/* The learned search's block (no reference counterpart): candidates given as
 * mode | slot << 3 | shape entry << 9 (at most 16).  Among the quality-0 fits
 * the block that decodes (reference decoder) to the smallest
 * ora_block_premul_err wins, ties to the earlier candidate; with quality > 0
 * the candidates within 1/50 of that error are annealed, the same choice is
 * made over annealed and remaining quality-0 fits, and the quality-0 winner
 * is kept if it decodes strictly better. */
void ora_bc7_encode_learned(const uint32_t px[16], const int32_t *cands, int n, uint32_t quality,
                            uint64_t seed, uint64_t block_id, uint8_t out[16]) {
  uint64_t best0 = UINT64_MAX, bestq = UINT64_MAX;
  uint8_t blk0[16], blkq[16];
  uint64_t err0[16];
  uint8_t cand0[16][16];
  memset(blk0, 0, 16);
  memset(blkq, 0, 16);
  for (int pass = 0; pass < (quality > 0 ? 2 : 1); pass++) {
    for (int i = 0; i < n && i < 16; i++) {
      const int mode = cands[i] & 7, slot = (cands[i] >> 3) & 63, entry = (cands[i] >> 9) & 1;
      uint8_t b[16];
      uint32_t dec[16];
      /* only candidates whose quality-0 block decodes within 1/50 of the
       * best are annealed; the others keep their quality-0 fit */
      if (pass == 1 && !(err0[i] <= best0 + best0 / 50)) {
        memcpy(b, cand0[i], 16);
      } else {
        encode_candidate(px, mode, slot, entry, pass ? quality : 0, seed, block_id, b);
      }
      decode_block(b, dec, 0);
      const uint64_t e = ora_block_premul_err(px, dec);
      if (pass == 0) {
        err0[i] = e;
        memcpy(cand0[i], b, 16);
      }
      if (pass == 0 && e < best0) {
        best0 = e;
        memcpy(blk0, b, 16);
      }
      if (pass == 1 && e < bestq) {
        bestq = e;
        memcpy(blkq, b, 16);
      }
    }
  }
  if (quality > 0 && !(best0 < bestq)) memcpy(out, blkq, 16);
  else memcpy(out, blk0, 16);
}

/* Exhaustive quality-0 search over every (mode, slot) the candidate ranker
 * scores (no reference counterpart; built from CompressionMode::Compress,
 * Compressor.cpp:1300-1371, one rotation / index mode at a time): slot =
 * shape for modes 0-3 and 7, rotation * 2 + index mode for mode 4, rotation
 * for mode 5, 0 for mode 6.  Each candidate is packed and decoded with the
 * reference decoder; returns mode * 64 + slot of the smallest decoded
 * premultiplied error (ora_block_premul_err); ties go to the earlier mode in
 * CompressClusters' order, then the lower slot. */
int ora_bc7_exhaustive_label(const uint32_t px[16], uint64_t *err_out) {
  static const int order[8] = {0, 2, 1, 3, 7, 4, 5, 6};
  static const int nslots[8] = {16, 64, 64, 64, 8, 4, 1, 64};
  uint64_t best = UINT64_MAX;
  int label = -1;
  for (int mi = 0; mi < 8; mi++) {
    const int mode = order[mi];
    for (int slot = 0; slot < nslots[mode]; slot++) {
      uint8_t blk[16];
      uint32_t dec[16];
      ora_bc7_encode_candidate(px, mode, slot, blk);
      decode_block(blk, dec, 0);
      const uint64_t e = ora_block_premul_err(px, dec);
      if (e < best) {
        best = e;
        label = mode * 64 + slot;
      }
    }
  }
  if (err_out) *err_out = best;
  return label;
}


/* The quality guard (not in the reference): with quality > 0, the block of
 * the same selection at quality 0 replaces the annealed one when it decodes
 * to a strictly smaller premultiplied error.  Returns 1 if it did. */
static int q0_guard(const shape_sel *sel, const uint32_t px[16], const ora_bc7_settings *s,
                    uint64_t block_id, uint8_t out[16]) {
  if (!s->q0_guard || s->quality == 0) return 0;
  ora_bc7_settings s0 = *s;
  s0.quality = 0;
  uint8_t out0[16];
  memset(out0, 0, 16);
  compress_clusters(sel, px, &s0, block_id, out0, NULL);
  if (memcmp(out0, out, 16) == 0) return 0;
  uint32_t d0[16], d1[16];
  decode_block(out0, d0, 0);
  decode_block(out, d1, 0);
  if (ora_block_premul_err(px, d0) < ora_block_premul_err(px, d1)) {
    memcpy(out, out0, 16);
    return 1;
  }
  return 0;
}


void ora_bc7_compress_block(const uint32_t px[16], uint8_t out[16], const ora_bc7_settings *s,
                            uint32_t wm_index, uint64_t block_id, ora_bc7_block_stats *st) {
  if (st) {
    st->path = 2;
    st->mode = -1;
    st->best_error = -1;
    for (int i = 0; i < 8; i++) st->mode_error[i] = -1.0;
  }
  memset(out, 0, 16);
  if (all_one_color(px)) {
    write_solid(px[0], out, wm_index);
    if (st) { st->path = 0; st->mode = 5; }
    return;
  }
  int transparent = 1;
  for (int i = 0; i < 16; i++)
    if ((float)((px[i] >> 24) & 0xFF) > 0.0f) {
      transparent = 0;
      break;
    }
  if (transparent) {
    write_transparent(out);
    if (st) { st->path = 1; st->mode = 6; }
    return;
  }
  shape_sel sel = box_selection(px, s->error_metric);
  sel.modes &= s->block_modes;
  compress_clusters(&sel, px, s, block_id, out, st);
  const int kept = q0_guard(&sel, px, s, block_id, out);
  if (st) st->q0_kept = kept;
}

/* BPTCC::Compress (Compressor.cpp:1473-1520) with GetBlock (1460-1466). */
void ora_bc7_compress_job(const uint8_t *in, uint8_t *out, uint32_t width, uint32_t height,
                          uint32_t x_start, uint32_t y_start, uint32_t x_end, uint32_t y_end,
                          const ora_bc7_settings *s, uint32_t *wm_counter) {
  const uint32_t *pix = (const uint32_t *)in;
  const uint32_t bw = width / 4;
  uint8_t *o = out + ((y_start / 4) * bw + x_start / 4) * 16;
  const uint32_t end_y = y_end < height - 4 ? y_end : height - 4;
  uint32_t sx = x_start;
  for (uint32_t j = y_start; j <= end_y; j += 4) {
    const uint32_t end_x = j == y_end ? x_end : width;
    for (uint32_t i = sx; i < end_x; i += 4) {
      uint32_t blk[16];
      for (int r = 0; r < 4; r++) memcpy(blk + 4 * r, pix + (j + r) * width + i, 16);
      uint32_t wm = 0;
      if (all_one_color(blk)) {
        *wm_counter = (*wm_counter + 1) % 9;
        wm = *wm_counter;
      }
      ora_bc7_compress_block(blk, o, s, wm, (uint64_t)(j / 4) * bw + i / 4, NULL);
      o += 16;
    }
    sx = 0;
  }
}

/* ------------------------------------------------------------------------ */
/* CompressWithStats (Compressor.cpp:1577-1623, 1862-2183).                 */
/* ------------------------------------------------------------------------ */

/* std::min(a, b): b only when strictly smaller. */
static double min_ref(double a, double b) { return b < a ? b : a; }

/* UpdateErrorEstimate (Compressor.cpp:1937-1944). */
static void update_estimate(double *est, int mode, double e) {
  if (est[mode] == -1.0 || e < est[mode]) est[mode] = e;
}

/* EstimateTwoClusterErrorStats (Compressor.cpp:1862-1898): QuantizedError
 * of the subset between its bounding-box corners, quantised as mode 1 (mask
 * 0xFFFCFCFC) and mode 3 (0xFFFEFEFE) endpoints; the estimates are the two
 * errors, the result 0.0001 + the smaller; 0 for a one-colour subset. */
static double est_two_stats(const clus *c, double se[2], v4 w) {
  const v4 v = v4_sub(c->mx, c->mn);
  if (v4_dot(v, v) == 0) {
    se[0] = se[1] = 0.0;
    return 0.0;
  }
  const double e1 = qerr(c, c->mn, c->mx, 8, 0xFFFCFCFCu, w, NULL, NULL);
  se[0] = e1 >= 0.0 ? e1 : min_ref(se[0], e1);
  const double e3 = qerr(c, c->mn, c->mx, 8, 0xFFFEFEFEu, w, NULL, NULL);
  se[1] = e3 >= 0.0 ? e3 : min_ref(se[1], e3);
  double error = 0.0001;
  error += min_ref(e1, e3);
  return error;
}

/* EstimateThreeClusterErrorStats (Compressor.cpp:1900-1935): as above with
 * mode 0 (0xFFF0F0F0) and mode 2 (0xFFF8F8F8) endpoints and four index
 * values; here the estimates already include the 0.0001. */
static double est_three_stats(const clus *c, double se[2], v4 w) {
  const v4 v = v4_sub(c->mx, c->mn);
  if (v4_dot(v, v) == 0) {
    se[0] = se[1] = 0.0;
    return 0.0;
  }
  const double e0 = 0.0001 + qerr(c, c->mn, c->mx, 4, 0xFFF0F0F0u, w, NULL, NULL);
  se[0] = e0 >= 0.0 ? e0 : min_ref(se[0], e0);
  const double e2 = 0.0001 + qerr(c, c->mn, c->mx, 4, 0xFFF8F8F8u, w, NULL, NULL);
  se[1] = e2 >= 0.0 ? e2 : min_ref(se[1], e2);
  double error = 0.0001;
  error += min_ref(e0, e2);
  return error;
}

/* The shape scan of the statistics block path (Compressor.cpp:2063-2176). */
static shape_sel stats_selection(const uint32_t px[16], ora_bc7_log_stats *ls,
                                 uint32_t metric) {
  const v4 w = metric_weights(metric);
  shape_sel r;
  r.modes = 0xFF;
  memset(r.sh_idx, 0, sizeof(r.sh_idx));
  memset(r.sh_parts, 0, sizeof(r.sh_parts));
  clus c;
  clus_init(&c, px);
  int opaque = 1;
  for (uint32_t i = 0; i < c.n; i++)
    if (fabs(c.pts[i].v[3] - 255.0f) > 1e-10) opaque = 0;
  /* mode 6's estimate: the whole block as one line (:2046-2061) */
  {
    const v4 v = v4_sub(c.mx, c.mn);
    if (v4_dot(v, v) == 0) {
      ls->mode_estimate[6] = 0.0;
    } else {
      const double e =
          0.0001 + qerr(&c, c.mn, c.mx, 4, 0xFEFEFEFEu, w, NULL, NULL);
      update_estimate(ls->mode_estimate, 6, e);
    }
  }
  double best[2] = {DBL_MAX, DBL_MAX};
  int path = 0;
  r.nshapes = 1;
  for (int i = 0; i < 64; i++) {
    c.shape = (uint32_t)i;
    c.nparts = 2;
    double err = 0.0, est[2] = {-1.0, -1.0};
    for (int ci = 0; ci < 2; ci++) {
      clus_set_partition(&c, (uint32_t)ci);
      double se[2] = {-1.0, -1.0};
      err += est_two_stats(&c, se, w);
      for (int ei = 0; ei < 2; ei++)
        if (se[ei] >= 0.0) est[ei] = est[ei] == -1.0 ? se[ei] : est[ei] + se[ei];
    }
    if (est[0] != -1.0) update_estimate(ls->mode_estimate, 1, est[0]);
    if (est[1] != -1.0) update_estimate(ls->mode_estimate, 3, est[1]);
    if (err < best[0]) {
      ls->two_shape[ls->num_two_shape++] = err;
      best[0] = err;
      r.sh_idx[0] = i;
      r.sh_parts[0] = 2;
    }
    if (err < 1e-9) {
      path = 2;
      r.modes = kTwoPartModes;
      break;
    }
  }
  /* the three-subset scan runs for opaque blocks even after the two-subset
   * scan stopped early */
  if (opaque) {
    r.nshapes++;
    for (int i = 0; i < 64; i++) {
      c.shape = (uint32_t)i;
      c.nparts = 3;
      double err = 0.0, est[2] = {-1.0, -1.0};
      for (int ci = 0; ci < 3; ci++) {
        clus_set_partition(&c, (uint32_t)ci);
        double se[2] = {-1.0, -1.0};
        err += est_three_stats(&c, se, w);
        for (int ei = 0; ei < 2; ei++)
          if (se[ei] >= 0.0) est[ei] = est[ei] == -1.0 ? se[ei] : est[ei] + se[ei];
      }
      if (est[0] != -1.0) update_estimate(ls->mode_estimate, 0, est[0]);
      if (est[1] != -1.0) update_estimate(ls->mode_estimate, 2, est[1]);
      if (err < best[1]) {
        ls->three_shape[ls->num_three_shape++] = err;
        best[1] = err;
        r.sh_idx[1] = i;
        r.sh_parts[1] = 3;
      }
      if (err < 1e-9) {
        path = 2;
        r.modes = kThreePartModes;
        break;
      }
    }
  }
  ls->path = path == 0 ? 3 : path;
  return r;
}

static int written_mode(const uint8_t blk[16]) {
  for (int m = 0; m < 8; m++)
    if ((blk[0] >> m) & 1) return m;
  return 8;
}

void ora_bc7_compress_block_logged(const uint32_t px[16], uint8_t out[16],
                                   const ora_bc7_settings *s, uint32_t wm_index,
                                   uint64_t block_id, ora_bc7_log_stats *ls) {
  memset(ls, 0, sizeof(*ls));
  for (int i = 0; i < 8; i++) ls->mode_estimate[i] = ls->mode_error[i] = -1.0;
  memset(out, 0, 16);
  if (all_one_color(px)) {
    write_solid(px[0], out, wm_index);
    ls->path = 0;
  } else {
    int transparent = 1;
    for (int i = 0; i < 16; i++)
      if ((float)((px[i] >> 24) & 0xFF) > 0.0f) transparent = 0;
    if (transparent) {
      write_transparent(out);
      ls->path = 1;
    } else {
      shape_sel sel = stats_selection(px, ls, s->error_metric);
      sel.modes &= s->block_modes;
      compress_clusters(&sel, px, s, block_id, out, NULL);
      q0_guard(&sel, px, s, block_id, out);
    }
  }
  ls->mode = written_mode(out);
}

void ora_bc7_compress_job_logged(const uint8_t *in, uint8_t *out, uint32_t width,
                                 uint32_t height, uint32_t x_start, uint32_t y_start,
                                 uint32_t x_end, uint32_t y_end, const ora_bc7_settings *s,
                                 uint32_t *wm_counter, ora_bc7_log_stats *stats) {
  const uint32_t *pix = (const uint32_t *)in;
  const uint32_t bw = width / 4;
  uint8_t *o = out + ((y_start / 4) * bw + x_start / 4) * 16;
  const uint32_t end_y = y_end < height - 4 ? y_end : height - 4;
  uint32_t sx = x_start;
  size_t k = 0;
  for (uint32_t j = y_start; j <= end_y; j += 4) {
    const uint32_t end_x = j == y_end ? x_end : width;
    for (uint32_t i = sx; i < end_x; i += 4) {
      uint32_t blk[16];
      for (int r = 0; r < 4; r++) memcpy(blk + 4 * r, pix + (j + r) * width + i, 16);
      uint32_t wm = 0;
      if (all_one_color(blk)) {
        *wm_counter = (*wm_counter + 1) % 9;
        wm = *wm_counter;
      }
      ora_bc7_compress_block_logged(blk, o, s, wm, (uint64_t)(j / 4) * bw + i / 4, &stats[k++]);
      o += 16;
    }
    sx = 0;
  }
}

/* ------------------------------------------------------------------------ */
/* A caller's shape selection (CompressionSettings::m_ShapeSelectionFn,     */
/* Compressor.cpp:1840-1858).                                               */
/* ------------------------------------------------------------------------ */

void ora_bc7_compress_job_selected(const uint8_t *in, uint8_t *out, uint32_t width,
                                   uint32_t height, uint32_t x_start, uint32_t y_start,
                                   uint32_t x_end, uint32_t y_end, const ora_bc7_settings *s,
                                   uint32_t *wm_counter, const ora_shape_selection *sels) {
  const uint32_t *pix = (const uint32_t *)in;
  const uint32_t bw = width / 4;
  uint8_t *o = out + ((y_start / 4) * bw + x_start / 4) * 16;
  const uint32_t end_y = y_end < height - 4 ? y_end : height - 4;
  uint32_t sx = x_start;
  size_t k = 0;
  for (uint32_t j = y_start; j <= end_y; j += 4) {
    const uint32_t end_x = j == y_end ? x_end : width;
    for (uint32_t i = sx; i < end_x; i += 4, k++) {
      uint32_t blk[16];
      for (int r = 0; r < 4; r++) memcpy(blk + 4 * r, pix + (j + r) * width + i, 16);
      memset(o, 0, 16);
      int transparent = 1;
      for (int q = 0; q < 16; q++)
        if ((float)((blk[q] >> 24) & 0xFF) > 0.0f) transparent = 0;
      if (all_one_color(blk)) {
        *wm_counter = (*wm_counter + 1) % 9;
        write_solid(blk[0], o, *wm_counter);
      } else if (transparent) {
        write_transparent(o);
      } else {
        const ora_shape_selection *g = &sels[k];
        shape_sel sel;
        memset(&sel, 0, sizeof(sel));
        sel.nshapes = g->num_shapes;
        for (int e = 0; e < 5; e++) {
          sel.sh_parts[e] = (int)g->shapes[e].num_partitions;
          sel.sh_idx[e] = (int)g->shapes[e].index;
        }
        sel.modes = g->selected_modes & s->block_modes;
        compress_clusters(&sel, blk, s, (uint64_t)(j / 4) * bw + i / 4, o, NULL);
        q0_guard(&sel, blk, s, (uint64_t)(j / 4) * bw + i / 4, o);
      }
      o += 16;
    }
    sx = 0;
  }
}

/* ------------------------------------------------------------------------ */
/* Decoder (Decompressor.cpp:32-367).                                       */
/* ------------------------------------------------------------------------ */

typedef struct {
  const uint8_t *p;
  int pos;
} bitr;
static uint32_t br_read(bitr *r, int n) {
  uint32_t v = 0;
  for (int i = 0; i < n; i++) {
    uint32_t b = (r->p[r->pos >> 3] >> (r->pos & 7)) & 1;
    r->pos++;
    v |= b << i;
  }
  return v;
}

typedef struct {
  int mode, rot, idx_mode, shape;
  uint8_t ep[3][2][4];
  uint8_t cidx[16], aidx[16];
} unpacked;

/* UnpackParams (Decompressor.cpp:32-190).  Returns 0 for the reserved
 * mode-8 encoding, which the reference does not handle. */
static int unpack(const uint8_t blk[16], unpacked *u) {
  bitr r = {blk, 0};
  int mode = 0;
  while (mode < 8 && !br_read(&r, 1)) mode++;
  if (mode >= 8) return 0;
  const mode_attr *a = &kModes[mode];
  const int ns = a->subsets;
  u->mode = mode;
  u->idx_mode = 0; u->rot = 0; u->shape = 0;
  if (ns > 1) u->shape = (int)br_read(&r, mode == 0 ? 4 : 6);
  else if (a->has_rot) {
    u->rot = (int)br_read(&r, 2);
    if (a->has_idx_mode) u->idx_mode = (int)br_read(&r, 1);
  }
  int cp = a->color_prec, ap = a->alpha_prec;
  for (int ch = 0; ch < 3; ch++)
    for (int i = 0; i < ns; i++)
      for (int e = 0; e < 2; e++) u->ep[i][e][ch] = (uint8_t)(br_read(&r, cp) << (8 - cp));
  for (int i = 0; i < ns; i++)
    for (int e = 0; e < 2; e++) u->ep[i][e][3] = ap == 0 ? 0xFF : (uint8_t)(br_read(&r, ap) << (8 - ap));
  if (a->pbit_type == PB_SHARED) {
    cp++; ap++;
    for (int i = 0; i < ns; i++) {
      uint32_t pb = br_read(&r, 1);
      for (int e = 0; e < 2; e++)
        for (int ch = 0; ch < 4; ch++) u->ep[i][e][ch] |= (uint8_t)(pb << (8 - (ch == 3 ? ap : cp)));
    }
  } else if (a->pbit_type == PB_NOTSHARED) {
    cp++; ap++;
    for (int i = 0; i < ns; i++)
      for (int e = 0; e < 2; e++) {
        uint32_t pb = br_read(&r, 1);
        for (int ch = 0; ch < 4; ch++) u->ep[i][e][ch] |= (uint8_t)(pb << (8 - (ch == 3 ? ap : cp)));
      }
  }
  for (int i = 0; i < ns; i++)
    for (int e = 0; e < 2; e++)
      for (int ch = 0; ch < 4; ch++) {
        const int prec = ch == 3 ? ap : cp;
        u->ep[i][e][ch] |= (uint8_t)(u->ep[i][e][ch] >> prec);
      }
  int prec = a->idx_bits;
  for (int i = 0; i < 16; i++) {
    const int s = subset_for(i, u->shape, ns);
    u->cidx[i] = (uint8_t)br_read(&r, anchor_for(s, u->shape, ns) == i ? prec - 1 : prec);
  }
  prec = a->alpha_idx_bits;
  if (prec == 0) {
    memcpy(u->aidx, u->cidx, 16);
  } else {
    for (int i = 0; i < 16; i++) {
      const int s = subset_for(i, u->shape, ns);
      u->aidx[i] = (uint8_t)br_read(&r, anchor_for(s, u->shape, ns) == i ? prec - 1 : prec);
    }
    if (u->idx_mode) {
      uint8_t t[16];
      memcpy(t, u->aidx, 16);
      memcpy(u->aidx, u->cidx, 16);
      memcpy(u->cidx, t, 16);
    }
  }
  return 1;
}

/* DecompressBC7Block (Decompressor.cpp:255-319).  spec != 0 selects the
 * index widths the BC7 specification assigns after the index-mode swap. */
static void decode_block(const uint8_t blk[16], uint32_t px[16], int spec) {
  unpacked u;
  if (!unpack(blk, &u)) {
    for (int i = 0; i < 16; i++) px[i] = 0;
    return;
  }
  const mode_attr *a = &kModes[u.mode];
  int nba = a->alpha_idx_bits, nbc = a->idx_bits;
  if (spec && u.idx_mode) {
    int t = nba; nba = nbc; nbc = t;
  }
  for (int i = 0; i < 16; i++) {
    const int s = subset_for(i, u.shape, a->subsets);
    uint32_t pixel = 0;
    for (int ch = 0; ch < 4; ch++) {
      uint32_t i0, i1, e1, e2;
      if (ch == 3 && nba > 0) {
        i0 = kInterp[nba - 1][u.aidx[i]][0];
        i1 = kInterp[nba - 1][u.aidx[i]][1];
      } else {
        i0 = kInterp[nbc - 1][u.cidx[i]][0];
        i1 = kInterp[nbc - 1][u.cidx[i]][1];
      }
      e1 = u.ep[s][0][ch];
      e2 = u.ep[s][1][ch];
      pixel |= ((((e1 * i0 + e2 * i1) + 32) >> 6) & 0xFF) << (8 * ch);
    }
    uint8_t *pb = (uint8_t *)&pixel;
    if (u.rot >= 1 && u.rot <= 3) {
      uint8_t t = pb[u.rot - 1];
      pb[u.rot - 1] = pb[3];
      pb[3] = t;
    }
    px[i] = pixel;
  }
}

void ora_bc7_decode_block_spec(const uint8_t blk[16], uint32_t px[16]) { decode_block(blk, px, 1); }

static void decompress_image(const uint8_t *in, uint8_t *out, uint32_t width, uint32_t height,
                             int spec) {
  uint32_t *o = (uint32_t *)out;
  for (uint32_t j = 0; j < height; j += 4)
    for (uint32_t i = 0; i < width; i += 4) {
      uint32_t px[16];
      decode_block(in, px, spec);
      const uint32_t w = width - i < 4 ? width - i : 4;
      const uint32_t h = height - j < 4 ? height - j : 4;
      for (uint32_t jj = 0; jj < h; jj++) memcpy(o + (j + jj) * width + i, px + 4 * jj, w * 4);
      in += 16;
    }
}

void ora_bc7_decompress(const uint8_t *in, uint8_t *out, uint32_t width, uint32_t height) {
  decompress_image(in, out, width, height, 0);
}

void ora_bc7_decompress_spec(const uint8_t *in, uint8_t *out, uint32_t width, uint32_t height) {
  decompress_image(in, out, width, height, 1);
}

/* ------------------------------------------------------------------------ */
/* PSNR (Base/src/Image.cpp:205-255).                                       */
/* ------------------------------------------------------------------------ */

double ora_premul_sse(const uint8_t *a, const uint8_t *b, uint32_t npix) {
  double mse = 0.0;
  for (uint32_t i = 0; i < npix; i++) {
    const uint8_t *p = a + 4 * i, *q = b + 4 * i;
    const double ra = (double)p[3] / 255.0, ua = (double)q[3] / 255.0;
    for (int c = 0; c < 3; c++) {
      double d = ra * ((double)p[c] * 1.0) - ua * ((double)q[c] * 1.0);
      mse += d * d;
    }
  }
  return mse;
}

uint64_t ora_block_premul_err(const uint32_t src[16], const uint32_t dec[16]) {
  uint64_t e = 0;
  for (int i = 0; i < 16; i++) {
    const int64_t a = (int64_t)(src[i] >> 24), b = (int64_t)(dec[i] >> 24);
    for (int c = 0; c < 3; c++) {
      const int64_t d = a * (int64_t)((src[i] >> (8 * c)) & 0xFF) -
                        b * (int64_t)((dec[i] >> (8 * c)) & 0xFF);
      e += (uint64_t)(d * d);
    }
  }
  return e;
}

double ora_psnr(const uint8_t *a, const uint8_t *b, uint32_t width, uint32_t height) {
  double mse = ora_premul_sse(a, b, width * height);
  mse /= (double)(width * height);
  const double maxi = 3.0 * 255.0 * 255.0;
  return 10 * log10(maxi / mse);
}
