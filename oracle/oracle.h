/*
 * CPU oracle for the per-4x4-block encode search of FasTC.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C, the
 * reference's CPU algorithms (BPTCEncoder/src/Compressor.cpp,
 * RGBAEndpoints.{h,cpp}, Decompressor.cpp, Base/src/Image.cpp) so that the
 * tests can check the HIP kernels against it.  Only tests/, the smoke() entry
 * and the benchmark's CPU-baseline leg may load it; the product library never
 * links or calls it.
 *
 * Every exported function cites the reference lines it follows.  Functions
 * whose results are not pinned by a golden vector, a closed form, an
 * invariant or a brute-force check would be marked "parity unpinned" in their
 * comment; none is.  The reference BC7 encoder cannot be built without its
 * CMake step, so its cluster fits are pinned stage by stage instead
 * (tests/test_oracle_fit.py): GetPrincipalAxis against the float64
 * eigenvector and hand-derived known answers of its collinear shortcut,
 * k-means at its fixed point, least squares against the float64 normal
 * equations, ClampEndpointsToGrid by float64 brute force, whole fits by
 * exact two-colour subsets, the rotation fit by known answers of its
 * stale-statistics and equal-alpha branches, BoxSelection by a rescan of all
 * shapes, and Pack by an independent reader of the BC7 specification.
 */
#ifndef FASTC_ORACLE_H_
#define FASTC_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Random-number source for simulated annealing.
 *  ORA_RNG_LCG     : the reference's process-global LCG
 *                    (Compressor.cpp:358-362), state in *lcg_state, which the
 *                    caller seeds (the reference seeds it with time(NULL)).
 *  ORA_RNG_COUNTER : a counter-based stream keyed by (seed, block, fit); the
 *                    same definition is used by the HIP kernels, so annealed
 *                    output is bit-comparable between CPU and GPU. */
enum { ORA_RNG_LCG = 0, ORA_RNG_COUNTER = 1 };

typedef struct {
  uint32_t quality;       /* m_NumSimulatedAnnealingSteps (BPTCCompressor.h:149) */
  uint32_t block_modes;   /* m_BlockModes mask (BPTCCompressor.h:139) */
  int      rng_kind;      /* ORA_RNG_* */
  uint32_t *lcg_state;    /* ORA_RNG_LCG: the shared g_seed */
  uint64_t seed;          /* ORA_RNG_COUNTER: job seed */
  /* Quality guard (not in the reference): with quality > 0, a block whose
   * annealed encoding decodes (reference decoder) to a larger premultiplied
   * RGB error than its quality-0 encoding is written as the quality-0 one. */
  int      q0_guard;
  uint32_t error_metric;  /* m_ErrorMetric (BPTCCompressor.h:73-77, 141-142):
                           * 0 uniform, 1 nonuniform */
} ora_bc7_settings;

/* Per-block statistics (mode chosen, path) for tests. */
typedef struct {
  int path;          /* 0 solid, 1 transparent, 2 searched */
  int mode;          /* mode written */
  double best_error; /* internal error of the chosen mode (searched path) */
  double mode_error[8]; /* min internal error per mode, -1 if not tried */
  int q0_kept;       /* the quality guard replaced the annealed encoding */
} ora_bc7_block_stats;

/* Exhaustive quality-0 search over all 285 (mode, slot) candidates (the
 * candidate ranker's outputs): returns mode * 64 + slot with the smallest
 * decoded premultiplied error (ora_block_premul_err), written to *err. */
int ora_bc7_exhaustive_label(const uint32_t px[16], uint64_t *err);

/* One (mode, slot) candidate at quality 0, packed into out; returns its
 * internal error. */
double ora_bc7_encode_candidate(const uint32_t px[16], int mode, int slot, uint8_t out[16]);

/* The learned search's choice among candidates (mode | slot << 3 |
 * entry << 9): smallest decoded premultiplied error, quality-0 fits kept when
 * annealing decodes worse (no reference counterpart). */
void ora_bc7_encode_learned(const uint32_t px[16], const int32_t *cands, int n, uint32_t quality,
                            uint64_t seed, uint64_t block_id, uint8_t out[16]);

/* Compress one 4x4 block (16 RGBA8 pixels, R in the low byte).
 * Follows CompressBC7Block (Compressor.cpp:1819-1860).
 * wm_index: the watermark table entry for a solid block (the reference's
 * gWMVal after its pre-increment, Compressor.cpp:1457).
 * block_id: global block index (keys the counter RNG). */
void ora_bc7_compress_block(const uint32_t px[16], uint8_t out[16],
                            const ora_bc7_settings *s, uint32_t wm_index,
                            uint64_t block_id, ora_bc7_block_stats *stats);

/* BPTCC::Compress over a (sub-)job (Compressor.cpp:1473-1520).
 * in: width*height RGBA8; out: base of the whole-image BC7 buffer.
 * *wm_counter is the reference's gWMVal (starts at 0xFFFFFFFF in a fresh
 * process); it is advanced once per solid block in processing order. */
void ora_bc7_compress_job(const uint8_t *in, uint8_t *out, uint32_t width,
                          uint32_t height, uint32_t x_start, uint32_t y_start,
                          uint32_t x_end, uint32_t y_end,
                          const ora_bc7_settings *s, uint32_t *wm_counter);

/* One block's statistics from BPTCC::CompressWithStats, in the vocabulary
 * of its log (EBlockStats, Compressor.cpp:77-131; the statistics version of
 * CompressBC7Block, :1954-2183).  Same layout as fh_bc7_block_stats.
 *  path            BlockStat_Path: 0 solid, 1 transparent, 2 a shape scan
 *                  stopped at an estimate below 1e-9, 3 otherwise
 *  mode            BlockStat_Mode: the mode of the block written
 *  num_two_shape   how many BlockStat_TwoShapeEstimate lines (one per strict
 *                  improvement of the two-subset scan), values in two_shape
 *  num_three_shape the same for BlockStat_ThreeShapeEstimate
 *  mode_estimate   BlockStat_Mode<N>Estimate, -1 when never set
 *  mode_error      BlockStat_Mode<N>Error: always -1, since CompressClusters
 *                  keeps std::min(error, -1.0) (Compressor.cpp:1800-1801) */
typedef struct {
  int32_t path, mode;
  uint32_t num_two_shape, num_three_shape;
  double mode_estimate[8];
  double mode_error[8];
  double two_shape[64];
  double three_shape[64];
} ora_bc7_log_stats;

/* The statistics version of CompressBC7Block (Compressor.cpp:1954-2183):
 * its own shape scan (EstimateTwo/ThreeClusterErrorStats, :1862-1935; the
 * opaque test is A == 255, modes 4 and 5 stay selected for opaque blocks),
 * then CompressClusters and, as in ora_bc7_compress_block, the optional
 * quality guard. */
void ora_bc7_compress_block_logged(const uint32_t px[16], uint8_t out[16],
                                   const ora_bc7_settings *s, uint32_t wm_index,
                                   uint64_t block_id, ora_bc7_log_stats *ls);

/* BPTCC::CompressWithStats over a (sub-)job (Compressor.cpp:1577-1623):
 * the loop of ora_bc7_compress_job; stats[k] receives the k-th block
 * processed. */
void ora_bc7_compress_job_logged(const uint8_t *in, uint8_t *out, uint32_t width,
                                 uint32_t height, uint32_t x_start, uint32_t y_start,
                                 uint32_t x_end, uint32_t y_end,
                                 const ora_bc7_settings *s, uint32_t *wm_counter,
                                 ora_bc7_log_stats *stats);

/* BPTCC::ShapeSelection (BPTCEncoder/include/FasTC/BPTCCompressor.h:85-112);
 * same layout as fh_shape_selection. */
typedef struct {
  uint32_t num_shapes;                                /* m_NumShapesToSearch */
  struct { uint32_t num_partitions, index; } shapes[10];  /* m_Shapes */
  uint32_t selected_modes;                            /* m_SelectedModes */
} ora_shape_selection;

/* BPTCC::Compress with a caller's shape selection function
 * (CompressionSettings::m_ShapeSelectionFn, Compressor.cpp:1840-1858):
 * sels[k] is the selection of the k-th block processed (ignored for solid
 * and transparent blocks, which never reach the selection function); the
 * selected modes are and-ed with block_modes, then CompressClusters runs over
 * the first min(5, num_shapes) entries. */
void ora_bc7_compress_job_selected(const uint8_t *in, uint8_t *out, uint32_t width,
                                   uint32_t height, uint32_t x_start, uint32_t y_start,
                                   uint32_t x_end, uint32_t y_end, const ora_bc7_settings *s,
                                   uint32_t *wm_counter, const ora_shape_selection *sels);

/* BPTCC::Decompress (Decompressor.cpp:345-367), including the reference's
 * treatment of mode 4 with index mode 1 (Decompressor.cpp:263-295 interpolate
 * with the per-mode bit counts, not the swapped ones). */
/* The whole image by the BC7 specification's decode (ora_bc7_decode_block_spec). */
void ora_bc7_decompress_spec(const uint8_t *in, uint8_t *out, uint32_t width, uint32_t height);
void ora_bc7_decompress(const uint8_t *in, uint8_t *out, uint32_t width,
                        uint32_t height);

/* Spec-correct BC7 decode of one block (BC7 format specification); differs
 * from ora_bc7_decompress only for mode 4 with index mode 1. */
void ora_bc7_decode_block_spec(const uint8_t blk[16], uint32_t px[16]);

/* Image::ComputePSNR (Base/src/Image.cpp:205-255): premultiplied-alpha RGB. */
double ora_psnr(const uint8_t *a, const uint8_t *b, uint32_t width,
                uint32_t height);

/* Sum over pixels and RGB channels of the squared premultiplied difference,
 * the numerator of ComputePSNR's mse (Image.cpp:244-248). */
double ora_premul_sse(const uint8_t *a, const uint8_t *b, uint32_t npix);

/* Premultiplied RGB error of one block, scaled by 255^2 so that it is an
 * exact integer: sum over pixels and RGB of (a*c - a'*c')^2.  Orders blocks
 * the way Image::ComputePSNR's per-pixel term does. */
uint64_t ora_block_premul_err(const uint32_t src[16], const uint32_t dec[16]);

/* Building blocks exported for the function-level tests. */
uint8_t ora_quantize_channel(uint8_t val, uint8_t mask, int pbit);
uint32_t ora_to_pixel(const float v[4], uint32_t mask, int pbit);
float ora_length(const float v[4]);
int ora_power_method(const float m[16], float eig[4], float *eigval);
double ora_quantized_error(const uint32_t px[16], const uint8_t *sel, int n,
                           const float p1[4], const float p2[4], int nbuckets,
                           uint32_t mask, const int *pbits, uint8_t *indices);
/* The same with per-channel error weights w (the uniform metric is all 1). */
double ora_quantized_error_w(const uint32_t px[16], const uint8_t *sel, int n,
                             const float p1[4], const float p2[4], int nbuckets,
                             uint32_t mask, const int *pbits, const float w[4],
                             uint8_t *indices);
void ora_bc7_box_selection(const uint32_t px[16], uint32_t *nshapes,
                           uint32_t *shape2, uint32_t *shape3,
                           uint32_t *modes);
void ora_bc7_box_selection_m(const uint32_t px[16], uint32_t metric, uint32_t *nshapes,
                             uint32_t *shape2, uint32_t *shape3, uint32_t *modes);
const uint8_t *ora_table_mode5(void);   /* [256][2] 7-bit endpoint pairs */
const uint8_t *ora_table_dxt6(void);    /* [256][2][3] */
/* Partition masks (Shapes.h:23-50), anchors (AnchorTables.h:22-45) and
 * interpolation weights (Compressor.cpp:159-168), copied out flat. */
void ora_table_partitions(uint16_t shape2[64], uint16_t shape3[128], uint8_t anchor2[64],
                          uint8_t anchor3[128], uint32_t interp[128]);
/* CompressionMode::CompressSingleColor (Compressor.cpp:252-353) for one
 * mode and index mode (rotation 0): endpoints p1/p2 (8-bit values as
 * floats), p-bit combination and the per-pixel error. */
/* Intermediate values of one RGB cluster fit (CompressCluster,
 * Compressor.cpp:921-1094), for the tests. */
enum { ORA_FIT_SAME_POINT = 1, ORA_FIT_ONE_BUCKET = 2, ORA_FIT_LEAST_SQUARES = 3 };
typedef struct {
  int path;               /* ORA_FIT_* */
  float single[4];        /* the colour given to CompressSingleColor (paths 1, 2) */
  float axis[4];          /* GetPrincipalAxis */
  float start1[4], start2[4];  /* avg + {min,max} dp * axis, clamped */
  uint32_t nbuckets;
  uint32_t iterations;    /* k-means passes until nothing moved */
  float centroid[16][4];  /* k-means centroids at the fixed point */
  uint32_t count[16];     /* points per bucket */
  uint8_t bucket[16];     /* bucket of each point, cluster order */
  float lsq1[4], lsq2[4]; /* least-squares endpoints (path 3) */
  float grid1[4], grid2[4];  /* after ClampEndpointsToGrid */
  int pbit_combo;
  /* rotation fits only (modes 4 and 5) */
  double rgb_error;        /* the rotated RGB fit's error */
  double alpha_error;      /* the alpha term */
  int alpha_path;          /* ORA_FIT_LEAST_SQUARES when the alpha values differ */
  float alpha_centroid[8]; /* 1-D k-means at its fixed point */
  uint32_t alpha_count[8];
  uint8_t alpha_bucket[16];
  float alpha_lsq[2];      /* least squares, clamped to [0, 255] */
  uint8_t alpha_q[2];      /* quantised after truncation to uint8 */
} ora_fit_trace;

/* GetPrincipalAxis of subset `subset` of shape `shape` (nparts subsets). */
void ora_principal_axis(const uint32_t px[16], int nparts, int shape, int subset, float axis[4]);
/* ClampEndpointsToGrid of `mode` (in place; combo index in *pbc). */
void ora_clamp_to_grid(int mode, float p1[4], float p2[4], int *pbc);
/* CompressCluster (RGB) of one subset at quality 0; trace optional.
 * idx receives the cluster-order indices. */
double ora_fit_rgb(const uint32_t px[16], int mode, int shape, int subset, int idx_mode,
                   ora_fit_trace *tr, float p1[4], float p2[4], uint8_t idx[16], int *pbc);
/* CompressCluster (alpha/rotation, modes 4 and 5) at quality 0; the trace
 * records the fit of the rotated RGB cluster. */
double ora_fit_alpha(const uint32_t px[16], int mode, int rot, int idx_mode, ora_fit_trace *tr,
                     float p1[4], float p2[4], uint8_t idx[16], uint8_t aidx[16]);
/* The two fits above under error metric `metric` (0 uniform, 1 nonuniform). */
double ora_fit_rgb_m(const uint32_t px[16], int mode, int shape, int subset, int idx_mode,
                     uint32_t metric, ora_fit_trace *tr, float p1[4], float p2[4],
                     uint8_t idx[16], int *pbc);
double ora_fit_alpha_m(const uint32_t px[16], int mode, int rot, int idx_mode, uint32_t metric,
                       ora_fit_trace *tr, float p1[4], float p2[4], uint8_t idx[16],
                       uint8_t aidx[16]);
/* CompressionMode::Pack of given parameters: ep[subset][endpoint][channel]
 * (unquantised floats, quantised by Pack), p-bit combos per subset, the
 * colour (idx) and alpha (aidx) index of every pixel. */
void ora_bc7_pack(int mode, int shape, int rot, int idx_mode, const float ep[3][2][4],
                  const uint8_t pbc[3], const uint8_t idx[16], const uint8_t aidx[16],
                  uint8_t out[16]);
double ora_single_color(int mode, int idx_mode, const float p[4], float p1[4], float p2[4],
                        int *pbc);
double ora_single_color_m(int mode, int idx_mode, uint32_t metric, const float p[4],
                          float p1[4], float p2[4], int *pbc);
uint32_t ora_counter_rng(uint32_t key, uint32_t n);
/* QuantizedError point evaluations made by the annealing loops
 * (OptimizeEndpointsForCluster, Compressor.cpp:583-624) on the calling
 * thread since the last reset: one per point of the annealed subset per
 * step.  reset != 0 clears the count after reading it. */
uint64_t ora_anneal_point_count(int reset);
uint32_t ora_fit_key(uint64_t seed, uint64_t block_id, uint32_t fit_tag);

#ifdef __cplusplus
}
#endif

#endif
