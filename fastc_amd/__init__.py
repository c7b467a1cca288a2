"""fastc_amd: MI355X block-texture encoders (BC7/BPTC, DXT1/DXT5, ETC1).

Python host bindings over the C ABI in include/fasthip.h (libfasthip.so,
built in-tree by fastc_amd/build.py).  The encoders run only as HIP kernels;
importing this module fails loudly if the library is missing.

The names mirror FasTC's API: `CompressImageData` (Core/src/TexComp.cpp:427),
`SCompressionSettings` (Core/include/FasTC/TexComp.h:30-76) and the
`ECompressionFormat` values (Base/include/FasTC/CompressionFormat.h:26-30).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfasthip.so")

# FasTC::ECompressionFormat
eCompressionFormat_DXT1 = 0
eCompressionFormat_DXT5 = 1
eCompressionFormat_ETC1 = 2
eCompressionFormat_BPTC = 3
# BPTCC::ErrorMetric (BPTCCompressor.h:73-77)
eErrorMetric_Uniform = 0
eErrorMetric_Nonuniform = 1
# or'ed into a BPTC format for decoding: the BC7 specification's decode
# (FH_DECODE_SPEC) instead of the reference decoder's
DECODE_SPEC = 0x100

BLOCK_BYTES = {0: 8, 1: 16, 2: 8, 3: 16}

FH_OK = 0
FH_FLAG_CALLER_WATERMARK = 1
FH_FLAG_NO_Q0_GUARD = 2
FH_FLAG_LEARNED = 4
STATUS_NAMES = {0: "FH_OK", 1: "FH_ERR_INVALID_ARGUMENT", 2: "FH_ERR_NO_DATA",
                3: "FH_ERR_BLOCK_DIMS", 4: "FH_ERR_UNKNOWN_FORMAT",
                5: "FH_ERR_BUFFER_TOO_SMALL", 6: "FH_ERR_UNSUPPORTED",
                7: "FH_ERR_NO_DEVICE", 8: "FH_ERR_HIP"}


class FhJob(ctypes.Structure):
    _fields_ = [("format", ctypes.c_uint32),
                ("in_", ctypes.c_void_p),
                ("out", ctypes.c_void_p),
                ("out_size", ctypes.c_size_t),
                ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("x_start", ctypes.c_uint32), ("y_start", ctypes.c_uint32),
                ("x_end", ctypes.c_uint32), ("y_end", ctypes.c_uint32),
                ("quality", ctypes.c_uint32),
                ("block_modes", ctypes.c_uint32),
                ("error_metric", ctypes.c_uint32),
                ("seed", ctypes.c_uint64),
                ("wm_counter", ctypes.POINTER(ctypes.c_uint32)),
                ("flags", ctypes.c_uint32),
                ("ranker", ctypes.c_void_p),
                ("ranker_topk", ctypes.c_uint32),
                ("ranker_threshold", ctypes.c_float)]


class FhTrainConfig(ctypes.Structure):
    _fields_ = [("epochs", ctypes.c_uint32), ("batch", ctypes.c_uint32), ("lr", ctypes.c_float),
                ("shuffle_seed", ctypes.c_uint64)]


class FhTrainReport(ctypes.Structure):
    _fields_ = [("first_loss", ctypes.c_double), ("epoch_loss", ctypes.c_double * 16),
                ("steps", ctypes.c_uint32), ("ms", ctypes.c_float)]


class FhSurface(ctypes.Structure):
    _fields_ = [("format", ctypes.c_uint32), ("in_", ctypes.c_void_p), ("out", ctypes.c_void_p),
                ("width", ctypes.c_uint32), ("height", ctypes.c_uint32)]


class FhBatchOptions(ctypes.Structure):
    _fields_ = [("quality", ctypes.c_uint32), ("block_modes", ctypes.c_uint32),
                ("error_metric", ctypes.c_uint32), ("seed", ctypes.c_uint64),
                ("flags", ctypes.c_uint32), ("ranker", ctypes.c_void_p),
                ("ranker_topk", ctypes.c_uint32), ("ranker_threshold", ctypes.c_float),
                ("wm_counter", ctypes.c_void_p)]


class FhBatchTune(ctypes.Structure):
    _fields_ = [("wg_per_cu", ctypes.c_uint32), ("anneal_unroll", ctypes.c_uint32),
                ("simple_wg", ctypes.c_uint32), ("best_ms", ctypes.c_float),
                ("trials", ctypes.c_uint32)]


class FastHipError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


_lib = None


def lib():
    """The loaded libfasthip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing; run fastc_amd/build.py "
                              "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.fh_compress.argtypes = [ctypes.POINTER(FhJob)]
        L.fh_compress.restype = ctypes.c_int
        L.fh_compress_device.argtypes = [ctypes.POINTER(FhJob), ctypes.c_void_p]
        L.fh_compress_device.restype = ctypes.c_int
        L.fh_compress_device_noalloc.argtypes = [ctypes.POINTER(FhJob), ctypes.c_void_p]
        L.fh_compress_device_noalloc.restype = ctypes.c_int
        L.fh_decompress.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_size_t,
                                    ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]
        L.fh_decompress.restype = ctypes.c_int
        L.fh_compressed_size.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.POINTER(ctypes.c_size_t)]
        L.fh_compressed_size.restype = ctypes.c_int
        L.fh_last_error.restype = ctypes.c_char_p
        L.fh_device_available.restype = ctypes.c_int
        L.fh_set_kernel_timing.argtypes = [ctypes.c_int]
        L.fh_set_kernel_timing.restype = ctypes.c_int
        L.fh_kernel_time.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_uint32)]
        L.fh_kernel_time.restype = ctypes.c_int
        L.fh_set_work_counting.argtypes = [ctypes.c_int]
        L.fh_set_work_counting.restype = ctypes.c_int
        L.fh_bc7_work_counts.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
        L.fh_bc7_work_counts.restype = ctypes.c_int
        L.fh_bc7_exhaustive_labels.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p]
        L.fh_bc7_exhaustive_labels.restype = ctypes.c_int
        L.fh_bc7_candidate_errors.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                              ctypes.c_void_p]
        L.fh_bc7_candidate_errors.restype = ctypes.c_int
        vp = ctypes.c_void_p
        L.fh_ranker_create.argtypes = [ctypes.POINTER(vp)]
        L.fh_ranker_destroy.argtypes = [vp]
        L.fh_ranker_num_params.restype = ctypes.c_size_t
        L.fh_ranker_init.argtypes = [vp, ctypes.c_uint64]
        L.fh_ranker_set_params.argtypes = [vp, vp, ctypes.c_size_t]
        L.fh_ranker_get_params.argtypes = [vp, vp, ctypes.c_size_t]
        L.fh_ranker_forward.argtypes = [vp, vp, ctypes.c_uint32, vp, vp]
        L.fh_ranker_gradients.argtypes = [vp, vp, vp, ctypes.c_uint32, vp,
                                          ctypes.POINTER(ctypes.c_double)]
        L.fh_ranker_train.argtypes = [vp, vp, vp, ctypes.c_uint32, ctypes.POINTER(FhTrainConfig),
                                      ctypes.POINTER(FhTrainReport)]
        L.fh_batch_create.argtypes = [ctypes.POINTER(FhSurface), ctypes.c_uint32,
                                      ctypes.POINTER(FhBatchOptions), ctypes.POINTER(vp)]
        L.fh_batch_destroy.argtypes = [vp]
        L.fh_batch_run.argtypes = [vp, ctypes.c_int, vp]
        L.fh_batch_autotune.argtypes = [vp, vp, ctypes.POINTER(FhBatchTune)]
        L.fh_batch_blocks.argtypes = [vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
        L.fh_batch_set_tune.argtypes = [vp, ctypes.POINTER(FhBatchTune)]
        u32 = ctypes.c_uint32
        L.fh_decompress_device.argtypes = [u32, vp, vp, u32, u32, vp]
        L.fh_block_error_device.argtypes = [u32, vp, vp, u32, u32, vp, vp, vp]
        L.fh_image_error_device.argtypes = [vp, vp, u32, u32, vp, vp, vp]
        L.fh_psnr.argtypes = [vp, vp, u32, u32, ctypes.POINTER(ctypes.c_double)]
        L.fh_psnr_from_sse.argtypes = [ctypes.POINTER(ctypes.c_ulonglong), u32, u32]
        L.fh_psnr_from_sse.restype = ctypes.c_double
        L.fh_mip_level_device.argtypes = [vp, u32, u32, vp, vp]
        L.fh_pad4_device.argtypes = [vp, u32, u32, vp, vp]
        L.fh_ktx_write.argtypes = [u32, u32, u32, u32, ctypes.POINTER(ctypes.c_void_p), vp,
                                   ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.fh_compress_with_stats.argtypes = [ctypes.POINTER(FhJob), vp, ctypes.c_size_t]
        L.fh_compress_with_selection.argtypes = [ctypes.POINTER(FhJob), vp, ctypes.c_size_t]
        L.fh_compress_with_selection_device.argtypes = [ctypes.POINTER(FhJob), vp,
                                                        ctypes.c_size_t, vp]
        L.fh_compress_with_stats_device.argtypes = [ctypes.POINTER(FhJob), vp, ctypes.c_size_t,
                                                    vp]
        L.fh_bc7_stats_format.argtypes = [vp, ctypes.c_uint64, vp, ctypes.c_size_t,
                                          ctypes.POINTER(ctypes.c_size_t)]
        for name in ("fh_decompress_device", "fh_block_error_device", "fh_image_error_device",
                     "fh_psnr", "fh_mip_level_device", "fh_pad4_device", "fh_ktx_write",
                     "fh_compress_with_stats", "fh_compress_with_stats_device",
                     "fh_bc7_stats_format", "fh_compress_with_selection",
                     "fh_compress_with_selection_device"):
            getattr(L, name).restype = ctypes.c_int
        for name in ("fh_batch_create", "fh_batch_destroy", "fh_batch_run", "fh_batch_autotune",
                     "fh_batch_blocks", "fh_batch_set_tune"):
            getattr(L, name).restype = ctypes.c_int
        for name in ("fh_ranker_create", "fh_ranker_destroy", "fh_ranker_init",
                     "fh_ranker_set_params", "fh_ranker_get_params", "fh_ranker_forward",
                     "fh_ranker_gradients", "fh_ranker_train"):
            getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


EXPORTED_SYMBOLS = ["fh_compress", "fh_compress_device", "fh_decompress",
                    "fh_compressed_size", "fh_last_error", "fh_device_available",
                    "fh_set_kernel_timing", "fh_kernel_time", "fh_set_work_counting",
                    "fh_bc7_work_counts", "fh_bc7_exhaustive_labels",
                    "fh_ranker_create", "fh_ranker_destroy", "fh_ranker_num_params",
                    "fh_ranker_init", "fh_ranker_set_params", "fh_ranker_get_params",
                    "fh_ranker_forward", "fh_ranker_gradients", "fh_ranker_train",
                    "fh_batch_create", "fh_batch_destroy", "fh_batch_run", "fh_batch_autotune",
                    "fh_batch_blocks", "fh_batch_set_tune", "fh_decompress_device",
                    "fh_block_error_device", "fh_image_error_device", "fh_psnr_from_sse",
                    "fh_psnr", "fh_mip_level_device", "fh_pad4_device", "fh_ktx_write",
                    "fh_compress_with_stats", "fh_compress_with_stats_device",
                    "fh_bc7_stats_format", "fh_compress_with_selection",
                    "fh_compress_with_selection_device", "fh_compress_device_noalloc",
                    "fh_bc7_candidate_errors"]

# fh_shape_selection (include/fasthip.h): BPTCC::ShapeSelection
SHAPE_SELECTION_DTYPE = np.dtype([("num_shapes", "<u4"),
                                  ("shapes", [("num_partitions", "<u4"), ("index", "<u4")],
                                   (10,)),
                                  ("selected_modes", "<u4")])

# fh_bc7_block_stats (include/fasthip.h)
BLOCK_STATS_DTYPE = np.dtype([("path", "<i4"), ("mode", "<i4"), ("num_two_shape", "<u4"),
                              ("num_three_shape", "<u4"), ("mode_estimate", "<f8", (8,)),
                              ("mode_error", "<f8", (8,)), ("two_shape", "<f8", (64,)),
                              ("three_shape", "<f8", (64,))])


def _check(status):
    if status != FH_OK:
        raise FastHipError(status, lib().fh_last_error().decode(errors="replace"))


def device_available():
    return bool(lib().fh_device_available())


def compressed_size(fmt, width, height):
    n = ctypes.c_size_t(0)
    _check(lib().fh_compressed_size(fmt, width, height, ctypes.byref(n)))
    return n.value


def _job(fmt, inp, out, out_size, width, height, rect, quality, seed, block_modes,
         wm_counter, q0_guard=True, ranker=None, topk=4, threshold=0.6, metric=0):
    x0, y0, x1, y1 = rect if rect is not None else (0, 0, width, height)
    j = FhJob()
    j.format = fmt
    j.in_ = inp
    j.out = out
    j.out_size = out_size
    j.width, j.height = width, height
    j.x_start, j.y_start, j.x_end, j.y_end = x0, y0, x1, y1
    j.quality = quality
    j.block_modes = block_modes
    j.error_metric = metric
    j.seed = seed
    j.flags = 0 if q0_guard else FH_FLAG_NO_Q0_GUARD
    if ranker is not None:
        j.flags |= FH_FLAG_LEARNED
        j.ranker = ranker.handle
        j.ranker_topk = topk
        j.ranker_threshold = threshold
    if wm_counter is not None:
        j.wm_counter = ctypes.pointer(wm_counter)
        j.flags |= FH_FLAG_CALLER_WATERMARK
    return j


def compress(img, fmt=eCompressionFormat_BPTC, quality=0, seed=0, block_modes=0xFF,
             rect=None, out=None, wm_counter=None, q0_guard=True, ranker=None, topk=4,
             threshold=0.6, metric=eErrorMetric_Uniform):
    """Encode a (h, w) uint32 RGBA8 image (R in the low byte) on the GPU.

    Mirrors CompressImageData: returns the compressed bytes (uint8 array of
    CompressedImage::GetCompressedSize bytes).  `wm_counter` is a
    ctypes.c_uint32 holding the BC7 solid-block watermark counter (the
    reference's gWMVal); it is updated in place.  Without it the library's
    process-wide counter is used, like the reference.  `q0_guard=False`
    drops the BPTC quality guard (FH_FLAG_NO_Q0_GUARD).  With a `Ranker`, BPTC
    runs the learned candidate pruning (FH_FLAG_LEARNED) with `topk` and
    `threshold`.  `metric` is BPTCC's m_ErrorMetric (eErrorMetric_Uniform or
    eErrorMetric_Nonuniform)."""
    img = np.ascontiguousarray(img, dtype=np.uint32)
    h, w = img.shape
    size = compressed_size(fmt, w, h)
    if out is None:
        out = np.zeros(size, dtype=np.uint8)
    j = _job(fmt, img.ctypes.data, out.ctypes.data, out.nbytes, w, h, rect, quality, seed,
             block_modes, wm_counter, q0_guard, ranker, topk, threshold, metric)
    _check(lib().fh_compress(ctypes.byref(j)))
    return out


def compress_with_stats(img, quality=0, seed=0, block_modes=0xFF, rect=None, wm_counter=None,
                        q0_guard=True):
    """BPTCC::CompressWithStats on the GPU (fh_compress_with_stats): returns
    (compressed bytes, statistics), the statistics a BLOCK_STATS_DTYPE array
    with one record per block the job processes, in processing order."""
    img = np.ascontiguousarray(img, dtype=np.uint32)
    h, w = img.shape
    out = np.zeros(compressed_size(eCompressionFormat_BPTC, w, h), dtype=np.uint8)
    stats = np.zeros((h // 4) * (w // 4), dtype=BLOCK_STATS_DTYPE)
    j = _job(eCompressionFormat_BPTC, img.ctypes.data, out.ctypes.data, out.nbytes, w, h, rect,
             quality, seed, block_modes, wm_counter, q0_guard, None, 4, 0.6)
    _check(lib().fh_compress_with_stats(ctypes.byref(j), stats.ctypes.data, stats.size))
    return out, stats


def compress_with_stats_device(in_ptr, out_ptr, out_size, stats_ptr, stats_count, width,
                               height, quality=0, seed=0, block_modes=0xFF, stream=0,
                               wm_counter=None, q0_guard=True):
    """fh_compress_with_stats_device: device pointers for the image, the
    blocks and `stats_count` BLOCK_STATS_DTYPE records; asynchronous on
    `stream`."""
    j = _job(eCompressionFormat_BPTC, in_ptr, out_ptr, out_size, width, height, None, quality,
             seed, block_modes, wm_counter, q0_guard, None, 4, 0.6)
    _check(lib().fh_compress_with_stats_device(ctypes.byref(j), ctypes.c_void_p(stats_ptr),
                                               stats_count, ctypes.c_void_p(stream)))


def compress_with_selection(img, selections, quality=0, seed=0, block_modes=0xFF, rect=None,
                            wm_counter=None, q0_guard=True, metric=0):
    """BPTCC::Compress with a caller's shape selection function
    (fh_compress_with_selection): `selections` is a SHAPE_SELECTION_DTYPE
    array, one entry per block the job processes, in processing order."""
    img = np.ascontiguousarray(img, dtype=np.uint32)
    h, w = img.shape
    sels = np.ascontiguousarray(selections, dtype=SHAPE_SELECTION_DTYPE)
    out = np.zeros(compressed_size(eCompressionFormat_BPTC, w, h), dtype=np.uint8)
    j = _job(eCompressionFormat_BPTC, img.ctypes.data, out.ctypes.data, out.nbytes, w, h, rect,
             quality, seed, block_modes, wm_counter, q0_guard, None, 4, 0.6, metric)
    _check(lib().fh_compress_with_selection(ctypes.byref(j), sels.ctypes.data, sels.size))
    return out


def compress_with_selection_device(in_ptr, out_ptr, out_size, sel_ptr, sel_count, width, height,
                                   quality=0, seed=0, block_modes=0xFF, stream=0,
                                   wm_counter=None, q0_guard=True):
    """fh_compress_with_selection_device: device pointers for the image, the
    blocks and `sel_count` SHAPE_SELECTION_DTYPE records; asynchronous on
    `stream`."""
    j = _job(eCompressionFormat_BPTC, in_ptr, out_ptr, out_size, width, height, None, quality,
             seed, block_modes, wm_counter, q0_guard, None, 4, 0.6)
    _check(lib().fh_compress_with_selection_device(ctypes.byref(j), ctypes.c_void_p(sel_ptr),
                                                   sel_count, ctypes.c_void_p(stream)))


def stats_format(stats, block_index):
    """The log text CompressWithStats writes for one block
    (fh_bc7_stats_format): `stats` one BLOCK_STATS_DTYPE record."""
    rec = np.ascontiguousarray(np.asarray(stats, dtype=BLOCK_STATS_DTYPE).reshape(1))
    size = ctypes.c_size_t(0)
    _check(lib().fh_bc7_stats_format(rec.ctypes.data, block_index, None, 0, ctypes.byref(size)))
    buf = ctypes.create_string_buffer(size.value + 1)
    _check(lib().fh_bc7_stats_format(rec.ctypes.data, block_index, buf, size.value + 1,
                                     ctypes.byref(size)))
    return buf.value.decode()


def compress_device(in_ptr, out_ptr, out_size, width, height, fmt=eCompressionFormat_BPTC,
                    quality=0, seed=0, block_modes=0xFF, stream=0, wm_counter=None,
                    q0_guard=True, ranker=None, topk=4, threshold=0.6,
                    metric=eErrorMetric_Uniform):
    """Encode device-resident data (raw device pointers, e.g. torch
    tensor.data_ptr()) asynchronously on `stream` (a hipStream_t as int)."""
    j = _job(fmt, in_ptr, out_ptr, out_size, width, height, None, quality, seed, block_modes,
             wm_counter, q0_guard, ranker, topk, threshold, metric)
    _check(lib().fh_compress_device(ctypes.byref(j), ctypes.c_void_p(stream)))


def set_work_counting(enable):
    """Count the BC7 annealing's work (fh_set_work_counting)."""
    _check(lib().fh_set_work_counting(1 if enable else 0))


def work_counts():
    """(reference QuantizedError point evaluations, GPU pixel scores) of the
    last BPTC encode on the current device (fh_bc7_work_counts)."""
    c = (ctypes.c_uint64 * 2)()
    _check(lib().fh_bc7_work_counts(c))
    return int(c[0]), int(c[1])


def set_kernel_timing(enable):
    """Bracket every library kernel with HIP events on its stream (clears the
    totals)."""
    _check(lib().fh_set_kernel_timing(1 if enable else 0))


def kernel_time(kernel):
    """(total milliseconds, launches) of one kernel since timing was enabled."""
    ms = ctypes.c_double(0)
    n = ctypes.c_uint32(0)
    _check(lib().fh_kernel_time(kernel.encode(), ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value


def candidate_errors_device(blocks_ptr, n, errors_ptr, stream=0):
    """Every candidate's decoded error for the ranker's labels
    (fh_bc7_candidate_errors): blocks n x 16 uint32, errors n x 512 float32."""
    _check(lib().fh_bc7_candidate_errors(ctypes.c_void_p(blocks_ptr), n,
                                         ctypes.c_void_p(errors_ptr), ctypes.c_void_p(stream)))


def exhaustive_labels_device(blocks_ptr, n, labels_ptr, errors_ptr=0, stream=0):
    """Training labels of the BC7 candidate ranker on device buffers
    (fh_bc7_exhaustive_labels): blocks n x 16 uint32, labels n int32,
    errors n uint64 (optional)."""
    _check(lib().fh_bc7_exhaustive_labels(ctypes.c_void_p(blocks_ptr), n,
                                          ctypes.c_void_p(labels_ptr),
                                          ctypes.c_void_p(errors_ptr or None),
                                          ctypes.c_void_p(stream)))


class Ranker:
    """The BC7 candidate ranker (fh_ranker_*): a 64-256-256-512 MLP on the
    GPU.  Device buffers are raw pointers (e.g. torch tensor.data_ptr())."""

    def __init__(self, seed=None):
        self._h = ctypes.c_void_p()
        _check(lib().fh_ranker_create(ctypes.byref(self._h)))
        self.num_params = lib().fh_ranker_num_params()
        if seed is not None:
            self.init(seed)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.fh_ranker_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def init(self, seed):
        _check(lib().fh_ranker_init(self._h, seed))

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        _check(lib().fh_ranker_set_params(self._h, p.ctypes.data, p.size))

    def get_params(self):
        p = np.zeros(self.num_params, dtype=np.float32)
        _check(lib().fh_ranker_get_params(self._h, p.ctypes.data, p.size))
        return p

    def forward(self, blocks_ptr, n, logits_ptr, stream=0):
        _check(lib().fh_ranker_forward(self._h, ctypes.c_void_p(blocks_ptr), n,
                                       ctypes.c_void_p(logits_ptr), ctypes.c_void_p(stream)))

    def gradients(self, blocks_ptr, labels_ptr, n):
        g = np.zeros(self.num_params, dtype=np.float32)
        loss = ctypes.c_double(0)
        _check(lib().fh_ranker_gradients(self._h, ctypes.c_void_p(blocks_ptr),
                                         ctypes.c_void_p(labels_ptr), n, g.ctypes.data,
                                         ctypes.byref(loss)))
        return loss.value, g

    def train(self, blocks_ptr, labels_ptr, n, epochs, batch, lr, shuffle_seed=0):
        cfg = FhTrainConfig(epochs, batch, lr, shuffle_seed)
        rep = FhTrainReport()
        _check(lib().fh_ranker_train(self._h, ctypes.c_void_p(blocks_ptr),
                                     ctypes.c_void_p(labels_ptr), n, ctypes.byref(cfg),
                                     ctypes.byref(rep)))
        return {"first_loss": rep.first_loss,
                "epoch_loss": list(rep.epoch_loss)[:min(epochs, 16)],
                "steps": rep.steps, "ms": rep.ms}


class Batch:
    """A planned batch of surfaces (fh_batch_*).  `surfaces` is a list of
    (format, in_ptr, out_ptr, width, height) with device pointers; BPTC
    settings as in compress()."""

    def __init__(self, surfaces, quality=0, seed=0, block_modes=0xFF, q0_guard=True,
                 ranker=None, topk=4, threshold=0.6, wm_counter_ptr=0,
                 metric=eErrorMetric_Uniform):
        arr = (FhSurface * max(1, len(surfaces)))()
        for i, (fmt, inp, out, w, h) in enumerate(surfaces):
            arr[i] = FhSurface(fmt, inp, out, w, h)
        opt = FhBatchOptions()
        opt.quality = quality
        opt.block_modes = block_modes
        opt.error_metric = metric
        opt.seed = seed
        opt.flags = 0 if q0_guard else FH_FLAG_NO_Q0_GUARD
        if ranker is not None:
            opt.flags |= FH_FLAG_LEARNED
            opt.ranker = ranker.handle
            opt.ranker_topk = topk
            opt.ranker_threshold = threshold
        opt.wm_counter = wm_counter_ptr or None
        self._ranker = ranker
        self._h = ctypes.c_void_p()
        _check(lib().fh_batch_create(arr, len(surfaces), ctypes.byref(opt), ctypes.byref(self._h)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.fh_batch_destroy(h)
            self._h = None

    def run(self, use_graph=False, stream=0):
        _check(lib().fh_batch_run(self._h, 1 if use_graph else 0, ctypes.c_void_p(stream)))

    def autotune(self, stream=0):
        t = FhBatchTune()
        _check(lib().fh_batch_autotune(self._h, ctypes.c_void_p(stream), ctypes.byref(t)))
        return {"wg_per_cu": t.wg_per_cu, "anneal_unroll": t.anneal_unroll,
                "simple_wg": t.simple_wg, "best_ms": t.best_ms, "trials": t.trials}

    def set_tune(self, tune):
        t = FhBatchTune(tune["wg_per_cu"], tune["anneal_unroll"], tune["simple_wg"], 0.0, 0)
        _check(lib().fh_batch_set_tune(self._h, ctypes.byref(t)))

    def blocks(self, fmt):
        n = ctypes.c_uint32(0)
        _check(lib().fh_batch_blocks(self._h, fmt, ctypes.byref(n)))
        return n.value


def decompress(data, width, height, fmt=eCompressionFormat_BPTC):
    """Decode on the GPU to a (h, w) uint32 image (BPTCC::Decompress,
    DXTC::DecompressDXT1/DXT5, ETCC::Decompress)."""
    data = np.ascontiguousarray(data, dtype=np.uint8)
    out = np.zeros((height, width), dtype=np.uint32)
    _check(lib().fh_decompress(fmt, data.ctypes.data, data.nbytes, out.ctypes.data, width,
                               height))
    return out


def decompress_device(in_ptr, out_ptr, width, height, fmt=eCompressionFormat_BPTC, stream=0):
    """Decode device data to a device RGBA8 image, asynchronously on `stream`."""
    _check(lib().fh_decompress_device(fmt, ctypes.c_void_p(in_ptr), ctypes.c_void_p(out_ptr),
                                      width, height, ctypes.c_void_p(stream)))


def block_error_device(orig_ptr, enc_ptr, width, height, sse_ptr, block_err_ptr=0,
                       fmt=eCompressionFormat_BPTC, stream=0):
    """Premultiplied error of an encoded device surface against the original
    (fh_block_error_device): per-block uint64 errors (optional) and the
    128-bit total as two uint64 words at sse_ptr."""
    _check(lib().fh_block_error_device(fmt, ctypes.c_void_p(orig_ptr), ctypes.c_void_p(enc_ptr),
                                       width, height, ctypes.c_void_p(block_err_ptr or None),
                                       ctypes.c_void_p(sse_ptr), ctypes.c_void_p(stream)))


def image_error_device(a_ptr, b_ptr, width, height, sse_ptr, block_err_ptr=0, stream=0):
    """The same comparison between two RGBA8 device images."""
    _check(lib().fh_image_error_device(ctypes.c_void_p(a_ptr), ctypes.c_void_p(b_ptr), width,
                                       height, ctypes.c_void_p(block_err_ptr or None),
                                       ctypes.c_void_p(sse_ptr), ctypes.c_void_p(stream)))


def psnr_from_sse(sse, width, height):
    """PSNR (dB) from the two-word total of the error functions."""
    words = (ctypes.c_ulonglong * 2)(int(sse[0]), int(sse[1]))
    return lib().fh_psnr_from_sse(words, width, height)


def psnr(a, b):
    """Image::ComputePSNR of two (h, w) uint32 RGBA8 images, on the GPU."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    if a.shape != b.shape:
        return -1.0
    out = ctypes.c_double(0)
    _check(lib().fh_psnr(a.ctypes.data, b.ctypes.data, a.shape[1], a.shape[0],
                         ctypes.byref(out)))
    return out.value


def mip_level_device(src_ptr, width, height, dst_ptr, stream=0):
    """Next mip level of a device RGBA8 image into dst (max(1, w/2) x max(1, h/2))."""
    _check(lib().fh_mip_level_device(ctypes.c_void_p(src_ptr), width, height,
                                     ctypes.c_void_p(dst_ptr), ctypes.c_void_p(stream)))


def pad4_device(src_ptr, width, height, dst_ptr, stream=0):
    """Copy a device image into a ceil4(w) x ceil4(h) one padded with transparent black."""
    _check(lib().fh_pad4_device(ctypes.c_void_p(src_ptr), width, height,
                                ctypes.c_void_p(dst_ptr), ctypes.c_void_p(stream)))


def ktx_write(fmt, width, height, levels):
    """KTX 1.1 file bytes of an encoded texture; levels is a list of per-level
    encoded block arrays (level 0 first)."""
    datas = [np.ascontiguousarray(l, dtype=np.uint8) for l in levels]
    ptrs = (ctypes.c_void_p * len(datas))(*[d.ctypes.data for d in datas])
    size = ctypes.c_size_t(0)
    _check(lib().fh_ktx_write(fmt, width, height, len(datas), ptrs, None, 0, ctypes.byref(size)))
    out = np.zeros(size.value, dtype=np.uint8)
    _check(lib().fh_ktx_write(fmt, width, height, len(datas), ptrs, out.ctypes.data, out.nbytes,
                              ctypes.byref(size)))
    return out.tobytes()
