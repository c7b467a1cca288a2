/*
 * fasthip: MI355X (gfx950) block-texture encoders behind a C ABI.
 *
 * Drop-in boundary for FasTC's per-block encode path.  Each entry point
 * replaces one reference interface:
 *
 *   fh_compress          CompressImageData (Core/include/FasTC/TexComp.h:81-88,
 *                        Core/src/TexComp.cpp:427-525) and the per-format
 *                        CompressionFunc it selects (Core/src/CompressionFuncs.h:28):
 *                        BPTCC::Compress (BPTCEncoder/include/FasTC/BPTCCompressor.h:166),
 *                        DXTC::CompressImageDXT1/DXT5 (DXTEncoder/include/FasTC/DXTCompressor.h:24-25),
 *                        ETCC::Compress_RG (ETCEncoder/include/FasTC/ETCCompressor.h:35).
 *                        The fh_job fields are FasTC::CompressionJob's
 *                        (Base/include/FasTC/CompressionJob.h:40-140) plus the
 *                        BPTCC::CompressionSettings fields (BPTCCompressor.h:123-158).
 *   fh_compress_device   the same on device-resident buffers, asynchronous on
 *                        a caller stream (no host copies; used by benchmarks).
 *   fh_decompress        BPTCC::Decompress (BPTCCompressor.h:216),
 *                        DXTC::DecompressDXT1/5, ETCC::Decompress
 *                        (via CompressedImage::DecompressImage,
 *                        Core/src/CompressedImage.cpp:86-118); the _device
 *                        variant on device buffers.
 *   fh_psnr, fh_block_error_device, fh_image_error_device
 *                        Image::ComputePSNR (Base/src/Image.cpp:205-255) and
 *                        its per-block terms, fused with decoding.
 *   fh_compress_with_stats[_device], fh_bc7_stats_format
 *                        BPTCC::CompressWithStats (BPTCCompressor.h:173,
 *                        BPTCEncoder/src/Compressor.cpp:1577-1623) and the
 *                        BlockLogger lines it writes (Compressor.cpp:1373-1390).
 *   fh_compress_with_selection[_device]
 *                        BPTCC::Compress with a caller's m_ShapeSelectionFn
 *                        (BPTCCompressor.h:123-134, Compressor.cpp:1840-1858).
 *   fh_compressed_size   CompressedImage::GetCompressedSize
 *                        (Core/src/CompressedImage.cpp:136-149).
 *   fh_last_error        the stderr message CompressImageData prints
 *                        (Core/src/TexComp.cpp:157-159) before returning false.
 *   fh_device_available  no reference counterpart (the reference is CPU only);
 *                        lets callers choose this backend.
 *   fh_set_kernel_timing, fh_kernel_time
 *                        per-kernel timing for benchmarks (no counterpart).
 *
 * All functions are thread-safe.  Buffers are owned by the caller; no pointer
 * is retained after a call returns.
 */
#ifndef FASTHIP_H_
#define FASTHIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  FH_ERR_NO_DATA, FH_ERR_BLOCK_DIMS, FH_ERR_UNKNOWN_FORMAT and
 * FH_ERR_BUFFER_TOO_SMALL are the failures CompressImageData reports
 * (Core/src/TexComp.cpp:454-496). */
typedef enum fh_status {
  FH_OK = 0,
  FH_ERR_INVALID_ARGUMENT = 1,
  FH_ERR_NO_DATA = 2,
  FH_ERR_BLOCK_DIMS = 3,
  FH_ERR_UNKNOWN_FORMAT = 4,
  FH_ERR_BUFFER_TOO_SMALL = 5,
  FH_ERR_UNSUPPORTED = 6,
  FH_ERR_NO_DEVICE = 7,
  FH_ERR_HIP = 8
} fh_status;

/* Values of FasTC::ECompressionFormat (Base/include/FasTC/CompressionFormat.h:26-30). */
enum {
  FH_FORMAT_DXT1 = 0,
  FH_FORMAT_DXT5 = 1,
  FH_FORMAT_ETC1 = 2,
  FH_FORMAT_BPTC = 3
};

/* Or'ed into the format of fh_decompress, fh_decompress_device and
 * fh_block_error_device with FH_FORMAT_BPTC: decode by the BC7 specification
 * instead of the reference decoder.  No counterpart in the reference: its
 * decoder (Decompressor.cpp:263-295) keeps the per-mode index widths for
 * mode 4 blocks with index mode 1, which the specification (and hardware)
 * swaps with the indices; every other block decodes the same. */
enum { FH_DECODE_SPEC = 0x100 };

/* The BC7 candidate ranker (see the ranker section below). */
typedef struct fh_ranker fh_ranker;

/* fh_job.flags */
enum {
  /* Keep the solid-block watermark counter in *wm_counter (caller-owned)
   * instead of the library's process-wide counter. */
  FH_FLAG_CALLER_WATERMARK = 1u << 0,
  /* BPTC with quality > 0: skip the quality guard.  By default a block whose
   * annealed encoding decodes to a larger (premultiplied-alpha RGB) error
   * than its quality-0 encoding is written as the quality-0 encoding, so no
   * block is worse than at quality 0; the reference has no such guard. */
  FH_FLAG_NO_Q0_GUARD = 1u << 1,
  /* BPTC: learned candidate pruning (no reference counterpart).  The ranker
   * scores every block's (mode, slot) candidates; blocks whose top softmax
   * probability reaches ranker_threshold search only their ranker_topk best
   * candidates, the others fall back to BoxSelection's candidates.  Among a
   * block's candidates the one that decodes (reference decoder) to the
   * smallest premultiplied RGB error is written, quality-0 fits included, so
   * no block is worse than its quality-0 encoding. */
  FH_FLAG_LEARNED = 1u << 2
};

/* One compression job: FasTC::CompressionJob plus BPTC settings. */
typedef struct fh_job {
  uint32_t format;        /* FH_FORMAT_* */
  const uint8_t *in;      /* width*height*4 bytes, RGBA8, R first */
  uint8_t *out;           /* the whole image's compressed buffer */
  size_t out_size;        /* bytes available at out */
  uint32_t width, height; /* pixels, multiples of 4 */
  /* Sub-rectangle in block-raster order (CompressionJob XStart..YEnd).
   * A full image is x_start=y_start=0, x_end=width, y_end=height. */
  uint32_t x_start, y_start, x_end, y_end;
  /* BPTC only (BPTCC::CompressionSettings) */
  uint32_t quality;       /* simulated-annealing steps; the reference sets it from iQuality */
  uint32_t block_modes;   /* mask of BC7 modes (bit m = mode m); 0 means all */
  uint32_t error_metric;  /* m_ErrorMetric: 0 uniform (tc's), 1 nonuniform (weights
                           * sqrt(0.3), sqrt(0.56), sqrt(0.11), 1; reference search
                           * only: not with FH_FLAG_LEARNED or statistics) */
  uint64_t seed;          /* annealing random stream key */
  uint32_t *wm_counter;   /* with FH_FLAG_CALLER_WATERMARK: in/out watermark counter */
  uint32_t flags;
  /* with FH_FLAG_LEARNED */
  const fh_ranker *ranker;
  uint32_t ranker_topk;     /* candidates per confident block, 1..8 (0 means 4) */
  float ranker_threshold;   /* top softmax probability needed to prune */
} fh_job;

fh_status fh_compress(const fh_job *job);
fh_status fh_compress_device(const fh_job *job, void *hip_stream);

/* No counterpart in the reference.  fh_compress_device for BPTC jobs that
 * allocates nothing, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The encode works in a fixed area kept in the library's code
 * object, 72 MiB per device, held until the process ends; jobs of at most
 * 16384 blocks (e.g. 512x512 texels; split larger images into sub-jobs).
 * Same output as fh_compress_device, the process-wide watermark counter
 * included.  Not with FH_FLAG_LEARNED or FH_FLAG_CALLER_WATERMARK
 * (FH_ERR_UNSUPPORTED), BPTC only.  Every call on a device shares that
 * device's area: calls that overlap on one device (other threads, other
 * streams not ordered by the caller) race. */
fh_status fh_compress_device_noalloc(const fh_job *job, void *hip_stream);

/* One block's statistics, in the vocabulary of the log CompressWithStats
 * writes (EBlockStats, BPTCEncoder/src/Compressor.cpp:77-131; the statistics
 * version of CompressBC7Block, :1954-2183). */
typedef struct fh_bc7_block_stats {
  int32_t path;             /* BlockStat_Path: 0 solid, 1 transparent, 2 a shape
                               scan stopped at an estimate below 1e-9, 3 otherwise */
  int32_t mode;             /* BlockStat_Mode: the mode of the block written */
  uint32_t num_two_shape;   /* BlockStat_TwoShapeEstimate lines: the two-subset
                               scan's strict improvements, in two_shape */
  uint32_t num_three_shape; /* BlockStat_ThreeShapeEstimate lines, in three_shape */
  double mode_estimate[8];  /* BlockStat_Mode<N>Estimate; -1 when never set */
  double mode_error[8];     /* BlockStat_Mode<N>Error: always -1, as in the
                               reference, whose CompressClusters keeps
                               std::min(error, -1.0) (Compressor.cpp:1800-1801) */
  double two_shape[64];
  double three_shape[64];
} fh_bc7_block_stats;

/* BPTCC::CompressWithStats (BPTCEncoder/include/FasTC/BPTCCompressor.h:173,
 * Compressor.cpp:1577-1623) for a BPTC job on host buffers: the same job
 * rules and output layout as fh_compress, with the statistics block path
 * (Compressor.cpp:1954-2183), whose shape scan differs from Compress's
 * BoxSelection: it estimates two-subset shapes with mode 1 and mode 3
 * endpoints and three-subset shapes with mode 0 and mode 2 endpoints, calls a
 * block opaque only if every alpha is 255, and keeps modes 4 and 5 for opaque
 * blocks (each single-subset mode then runs once per shape entry).  So the
 * bytes differ from fh_compress's.  stats[k] receives the k-th block the job
 * processes (raster order from (x_start, y_start)); stats_count must cover
 * them.  Rows follow fh_compress's rule (Compressor.cpp:1478-1482): the
 * reference's statistics loop (:1583) lacks the min(YEnd, Height - 4) clamp
 * and would run one row past a whole-image job, out of bounds.
 * FH_FLAG_LEARNED is rejected; other formats return
 * FH_ERR_UNSUPPORTED, as the reference only collects BPTC statistics
 * (Core/src/TexComp.cpp:87-108). */
fh_status fh_compress_with_stats(const fh_job *job, fh_bc7_block_stats *stats,
                                 size_t stats_count);
/* The same on device buffers (job in/out and stats in device memory, the
 * records 8-byte aligned), asynchronous on hip_stream, as fh_compress_device. */
fh_status fh_compress_with_stats_device(const fh_job *job, fh_bc7_block_stats *stats,
                                        size_t stats_count, void *hip_stream);

/* BPTCC::ShapeSelection (BPTCEncoder/include/FasTC/BPTCCompressor.h:85-112),
 * field for field: the shapes CompressClusters tries for one block and the
 * modes it allows. */
typedef struct fh_shape_selection {
  uint32_t num_shapes;      /* m_NumShapesToSearch; CompressClusters tries at most 5 */
  struct {
    uint32_t num_partitions;  /* Shape::m_NumPartitions */
    uint32_t index;           /* Shape::m_Index, < 64 when num_partitions is 2 or 3 */
  } shapes[10];
  uint32_t selected_modes;  /* m_SelectedModes, and-ed with fh_job.block_modes */
} fh_shape_selection;

/* BPTCC::Compress with CompressionSettings::m_ShapeSelectionFn set
 * (BPTCCompressor.h:123-134, Compressor.cpp:1840-1858) for a BPTC job on host
 * buffers: selections[k] is the ShapeSelection the caller's function returns
 * for the k-th block the job processes (raster order from (x_start,
 * y_start)); CompressClusters, annealing and packing run on the GPU.  Solid
 * and transparent blocks never reach the selection function; their entries
 * are ignored.  Every entry must describe at most 31 candidates (a rotation
 * mode counts 8 or 4, and with quality > 0 a single-subset mode counts once
 * per shape entry), else FH_ERR_UNSUPPORTED; an index of 64 or more on a
 * two- or three-partition shape is FH_ERR_INVALID_ARGUMENT (the reference
 * would read past its partition tables). */
fh_status fh_compress_with_selection(const fh_job *job, const fh_shape_selection *selections,
                                     size_t selections_count);
/* The same on device buffers (job in/out and the selections in device
 * memory, 4-byte aligned), asynchronous on hip_stream.  The selections are
 * not checked on the host here: a selection beyond 31 candidates keeps its
 * first 31 (in CompressClusters' order), and a shape index is taken modulo
 * 64, so no selection can make the kernels read outside their tables. */
fh_status fh_compress_with_selection_device(const fh_job *job,
                                            const fh_shape_selection *selections,
                                            size_t selections_count, void *hip_stream);

/* The text BlockLogger writes for one block (Compressor.cpp:1373-1390,
 * 1945-1950): a line "<block_index>: BlockStat_<name> -- <value>" per
 * statistic, in the order CompressWithStats writes them (shape estimates,
 * Path, Mode, then Mode<N>Estimate and Mode<N>Error for N = 0..7), doubles in
 * the ostream default format.  block_index is the block's raster index in the
 * image (CompressionJob::CoordsToBlockIdx).  Writes the text and a NUL when
 * out is non-null and large enough; *out_size receives the text's length
 * (without the NUL).  With out null only the size is computed. */
fh_status fh_bc7_stats_format(const fh_bc7_block_stats *stats, uint64_t block_index, char *out,
                              size_t out_capacity, size_t *out_size);

/* Decodes a width x height surface of any format to RGBA8 (host buffers:
 * in holds ceil(w/4) * ceil(h/4) blocks, out w * h * 4 bytes), bit-exact
 * with BPTCC::Decompress (BPTCEncoder/src/Decompressor.cpp:345-367),
 * DXTC::DecompressDXT1/DXT5 (DXTEncoder/src/Decompressor.cpp:96-155) and
 * ETCC::Decompress (ETCEncoder/src/Decompressor.cpp:27-48): DXT1 output
 * alpha is always 255, ETC1 alpha 255, edge blocks clipped to the image. */
fh_status fh_decompress(uint32_t format, const uint8_t *in, size_t in_size, uint8_t *out,
                        uint32_t width, uint32_t height);
/* The same on device buffers (in aligned to the block size, out to 4 bytes),
 * asynchronous on hip_stream. */
fh_status fh_decompress_device(uint32_t format, const uint8_t *in, uint8_t *out, uint32_t width,
                               uint32_t height, void *hip_stream);

/* Image comparison of Image::ComputePSNR (Base/src/Image.cpp:205-255).  The
 * error of a pixel pair is the sum over R, G, B of (a*c - a'*c')^2 with
 * 8-bit values, i.e. 255^2 times the reference's squared difference of
 * alpha-premultiplied channels; it is an exact integer, so the totals below
 * do not depend on summation order.
 *
 * fh_block_error_device decodes `encoded` (a surface of `format`) block by
 * block and compares it with `original` (RGBA8, width * height) without
 * writing the decoded image; fh_image_error_device compares two RGBA8
 * images.  Device buffers, blocks aligned to their size, pixels to 4 bytes,
 * error words to 8 bytes.  block_errors (NULL or ceil(w/4) * ceil(h/4)
 * words, raster block order) receives each block's error over its pixels
 * inside the image; sse[0] (low) and sse[1] (high) receive the 128-bit total.
 * Both are overwritten.  Asynchronous on hip_stream. */
fh_status fh_block_error_device(uint32_t format, const uint32_t *original, const uint8_t *encoded,
                                uint32_t width, uint32_t height,
                                unsigned long long *block_errors, unsigned long long *sse,
                                void *hip_stream);
fh_status fh_image_error_device(const uint32_t *a, const uint32_t *b, uint32_t width,
                                uint32_t height, unsigned long long *block_errors,
                                unsigned long long *sse, void *hip_stream);
/* PSNR in dB from such a total (host sse[2]): 10 log10(3 * 255^2 / mse) with
 * mse = sse / 255^2 / (width * height), +inf for identical images. */
double fh_psnr_from_sse(const unsigned long long sse[2], uint32_t width, uint32_t height);
/* Image::ComputePSNR of two RGBA8 host images of the same size. */
fh_status fh_psnr(const uint8_t *a, const uint8_t *b, uint32_t width, uint32_t height,
                  double *psnr);

/* Mip chains and containers (the reference encodes the levels it is given
 * and writes one level).
 *
 * fh_mip_level_device: the next mip level of a width x height RGBA8 image,
 * max(1, width/2) x max(1, height/2) texels (sides halve rounding down), each
 * the per-channel rounded mean of its 2x2 source texels (2x1 or 1x2 once a
 * side is 1): (a + b + c + d + 2) >> 2, (a + b + 1) >> 1.  No counterpart in
 * the reference.  fh_pad4_device: copies width x height texels into a
 * ceil4(width) x ceil4(height) image padded with transparent black, as
 * CompressImage pads (Core/src/TexComp.cpp:382-393).  Device buffers,
 * asynchronous on hip_stream. */
fh_status fh_mip_level_device(const uint32_t *src, uint32_t width, uint32_t height, uint32_t *dst,
                              void *hip_stream);
fh_status fh_pad4_device(const uint32_t *src, uint32_t width, uint32_t height, uint32_t *dst,
                         void *hip_stream);
/* KTX 1.1 file image of an encoded texture: ImageWriterKTX::WriteImage
 * (IO/src/ImageWriterKTX.cpp:69-160) byte for byte for one level of BPTC,
 * DXT1 or DXT5 whose sides are multiples of 4; levels > 1 writes a mip chain
 * (level i is max(1, width >> i) x max(1, height >> i) texels, level_data[i]
 * its grid of whole blocks) and ETC1 is written as GL_ETC1_RGB8_OES, both
 * beyond the reference.  Host buffers.  *out_size receives the file size;
 * with out == NULL nothing is written (size query). */
fh_status fh_ktx_write(uint32_t format, uint32_t width, uint32_t height, uint32_t levels,
                       const uint8_t *const *level_data, uint8_t *out, size_t out_capacity,
                       size_t *out_size);

fh_status fh_compressed_size(uint32_t format, uint32_t width, uint32_t height, size_t *size);

const char *fh_last_error(void);
int fh_device_available(void);

/* Training labels of the BC7 candidate ranker (no reference counterpart).
 * Device buffers: blocks = n x 16 RGBA8 pixels (block-major, raster order
 * inside a block), labels = n int32, errors = n uint64 or NULL.  For every
 * block, all 285 (mode, slot) candidates are fitted at quality 0 (slot =
 * partition shape for modes 0-3 and 7, rotation * 2 + index mode for mode 4,
 * rotation for mode 5, 0 for mode 6), packed and decoded with the reference
 * decoder; labels[i] = mode * 64 + slot of the smallest premultiplied RGB
 * error (sum of (a*c - a'*c')^2, written to errors[i]).  Ties go to the mode
 * BPTCC::Compress tries first (0,2,1,3,7,4,5,6), then the lower slot.
 * Asynchronous on hip_stream. */
fh_status fh_bc7_exhaustive_labels(const uint32_t *blocks, uint32_t n, int32_t *labels,
                                   unsigned long long *errors, void *hip_stream);
/* The same search's per-candidate errors (no reference counterpart):
 * errors = n x 512 floats on the device, errors[i * 512 + mode * 64 + slot]
 * the decoded premultiplied error of that candidate (float32), +inf for the
 * 227 logits that are no candidate; n < 2^23.  For label definitions other
 * than the arg-min (sets of near-best candidates).  Asynchronous. */
fh_status fh_bc7_candidate_errors(const uint32_t *blocks, uint32_t n, float *errors,
                                  void *hip_stream);

/* ---- BC7 candidate ranker (no reference counterpart) ----------------------
 * A 64 -> 256 -> 256 -> 512 MLP (ReLU, bf16 weights on the matrix cores,
 * fp32 accumulation, fp32 master weights) that scores the 512 (mode, slot)
 * logits of a block (logit = mode * 64 + slot, the numbering of
 * fh_bc7_exhaustive_labels; 285 are real candidates).  Inputs are the 16
 * pixels' RGBA / 255.  Parameters are one float array of
 * fh_ranker_num_params() = 214016 values:
 *   W1[64][256] b1[256] W2[256][256] b2[256] W3[256][512] b3[512]
 * (W[in][out]).  A ranker belongs to the device current at creation. */

typedef struct fh_train_config {
  uint32_t epochs;        /* passes over the training set */
  uint32_t batch;         /* blocks per Adam step */
  float lr;               /* Adam step size (beta1 0.9, beta2 0.999, eps 1e-8) */
  uint64_t shuffle_seed;  /* seed of the per-epoch permutations */
} fh_train_config;

typedef struct fh_train_report {
  double first_loss;      /* mean cross-entropy of the first batch */
  double epoch_loss[16];  /* mean cross-entropy of the last batch of each epoch */
  uint32_t steps;         /* Adam steps taken */
  float ms;               /* device time of the whole training */
} fh_train_report;

fh_status fh_ranker_create(fh_ranker **ranker);
fh_status fh_ranker_destroy(fh_ranker *ranker);
size_t fh_ranker_num_params(void);
/* He-uniform weights from a seed, zero biases, Adam state cleared. */
fh_status fh_ranker_init(fh_ranker *ranker, uint64_t seed);
fh_status fh_ranker_set_params(fh_ranker *ranker, const float *params, size_t count);
fh_status fh_ranker_get_params(fh_ranker *ranker, float *params, size_t count);
/* Device buffers: blocks n x 16 pixels, logits n x 512 floats. */
fh_status fh_ranker_forward(fh_ranker *ranker, const uint32_t *blocks, uint32_t n, float *logits,
                            void *hip_stream);
/* Mean softmax cross-entropy over the 285 real logits of one batch and its
 * gradient (host array of fh_ranker_num_params() floats); device inputs. */
fh_status fh_ranker_gradients(fh_ranker *ranker, const uint32_t *blocks, const int32_t *labels,
                              uint32_t n, float *grad, double *loss);
/* Trains on device-resident blocks and labels: cfg->epochs passes over a
 * fresh permutation each, one Adam step per cfg->batch blocks. */
fh_status fh_ranker_train(fh_ranker *ranker, const uint32_t *blocks, const int32_t *labels,
                          uint32_t n, const fh_train_config *cfg, fh_train_report *report);

/* ---- Batches (no single reference counterpart) ------------------------------
 * Many surfaces (textures and their mip levels) of mixed formats encoded by
 * one launch sequence: the GPU form of the reference's atomic job-list mode
 * (CompressionJobList, Base/include/FasTC/CompressionJob.h:172-209, driven
 * by CompressAtomic, BPTCEncoder/src/Compressor.cpp:1542-1574).  All blocks
 * of a format are numbered in submission order and persistent kernels take
 * them from device counters.  Each surface's output equals fh_compress_device
 * of that surface alone, with BPTC solid blocks taking watermark entries in
 * submission order from the process-wide counter, and annealing streams
 * keyed by the block's index among the batch's BPTC blocks.
 * All buffers are device memory; width and height are multiples of 4
 * (callers pad small mip levels as CompressImage does,
 * Core/src/TexComp.cpp:382-393). */
typedef struct fh_surface {
  uint32_t format;        /* FH_FORMAT_* */
  const uint8_t *in;      /* width * height * 4 bytes */
  uint8_t *out;           /* fh_compressed_size(format, width, height) bytes */
  uint32_t width, height;
} fh_surface;

/* BPTC settings of a batch: the fields of fh_job with the same names; the
 * watermark counter, if the caller keeps one, is a device word here. */
typedef struct fh_batch_options {
  uint32_t quality;
  uint32_t block_modes;
  uint32_t error_metric;
  uint64_t seed;
  uint32_t flags;           /* FH_FLAG_NO_Q0_GUARD, FH_FLAG_LEARNED */
  const fh_ranker *ranker;
  uint32_t ranker_topk;
  float ranker_threshold;
  uint32_t *wm_counter;     /* device word holding the watermark counter, or NULL for
                               the library's process-wide one */
} fh_batch_options;

/* Launch shapes chosen by fh_batch_autotune. */
typedef struct fh_batch_tune {
  uint32_t wg_per_cu;      /* persistent workgroups per CU (BPTC set-up and annealing) */
  uint32_t anneal_unroll;  /* annealing steps between refills of finished lanes */
  uint32_t simple_wg;      /* blocks per workgroup of the DXT / ETC1 kernels */
  float best_ms;           /* time of one run with these shapes */
  uint32_t trials;
} fh_batch_tune;

typedef struct fh_batch fh_batch;

/* Plans a batch: copies the surface table to the device and allocates every
 * work buffer, so that runs allocate nothing.  The surfaces' buffers must
 * stay valid while the plan is used. */
fh_status fh_batch_create(const fh_surface *surfaces, uint32_t count,
                          const fh_batch_options *options, fh_batch **batch);
fh_status fh_batch_destroy(fh_batch *batch);
/* Encodes every surface, asynchronously on hip_stream.  use_graph != 0
 * captures the launch sequence into a HIP graph on the first call and
 * replays it afterwards. */
fh_status fh_batch_run(fh_batch *batch, int use_graph, void *hip_stream);
/* Times the plan under a few launch shapes (coordinate descent) and keeps the
 * fastest; runs the batch several times (its outputs are overwritten). */
fh_status fh_batch_autotune(fh_batch *batch, void *hip_stream, fh_batch_tune *report);
/* Applies launch shapes found for another batch (wg_per_cu 8/16/32,
 * anneal_unroll 1/2/4, simple_wg 64..512). */
fh_status fh_batch_set_tune(fh_batch *batch, const fh_batch_tune *tune);
/* Number of blocks of one format in the batch. */
fh_status fh_batch_blocks(const fh_batch *batch, uint32_t format, uint32_t *blocks);

/* Kernel timing for benchmarks; no reference counterpart (the reference's
 * CompressImageData reports only a whole-job wall time, TexComp.cpp:517).
 * While enabled, every kernel the library launches is bracketed by HIP
 * events on its own stream.  Enabling or disabling clears the totals.
 * fh_kernel_time waits for the recorded events and returns the summed time
 * and launch count of one kernel: "bc7_scan", "bc7_select", "bc7_setup",
 * "bc7_anneal", "bc7_pack", "bc7_guard", "dxt", "etc1", "bc7_labels",
 * "bc7_rank", "decode", "block_error". */
fh_status fh_set_kernel_timing(int enable);
fh_status fh_kernel_time(const char *kernel, double *total_ms, uint32_t *launches);

/* Work counting for the BC7 annealing; no reference counterpart (a
 * diagnostic of the annealing's useful work, for the VALU roofline).  While
 * enabled, each BPTC encode (fh_compress, fh_compress_device, the selection
 * entry points; not the no-allocation entry point or the nonuniform metric)
 * clears and then fills two counters of its device: counts[0] the
 * QuantizedError point evaluations the reference's annealing makes for the
 * same encode (every step of a subset scores that subset's points,
 * Compressor.cpp:583-624), counts[1] the pixel scores the GPU's annealing
 * makes (16 per step of a candidate).  fh_bc7_work_counts waits for the
 * device and returns the counters of the calling thread's current device. */
fh_status fh_set_work_counting(int enable);
fh_status fh_bc7_work_counts(uint64_t counts[2]);

#ifdef __cplusplus
}
#endif

#endif /* FASTHIP_H_ */
