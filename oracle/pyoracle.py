"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's CPU-baseline leg.  The product never uses it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libfastc_ref.so")

RNG_LCG = 0
RNG_COUNTER = 1


class Settings(ctypes.Structure):
    _fields_ = [("quality", ctypes.c_uint32),
                ("block_modes", ctypes.c_uint32),
                ("rng_kind", ctypes.c_int),
                ("lcg_state", ctypes.POINTER(ctypes.c_uint32)),
                ("seed", ctypes.c_uint64),
                ("q0_guard", ctypes.c_int),
                ("error_metric", ctypes.c_uint32)]


class BlockStats(ctypes.Structure):
    _fields_ = [("path", ctypes.c_int), ("mode", ctypes.c_int),
                ("best_error", ctypes.c_double), ("mode_error", ctypes.c_double * 8),
                ("q0_kept", ctypes.c_int)]


class FitTrace(ctypes.Structure):
    """ora_fit_trace: intermediate values of one RGB cluster fit."""
    _fields_ = [("path", ctypes.c_int), ("single", ctypes.c_float * 4),
                ("axis", ctypes.c_float * 4), ("start1", ctypes.c_float * 4),
                ("start2", ctypes.c_float * 4), ("nbuckets", ctypes.c_uint32),
                ("iterations", ctypes.c_uint32), ("centroid", (ctypes.c_float * 4) * 16),
                ("count", ctypes.c_uint32 * 16), ("bucket", ctypes.c_uint8 * 16),
                ("lsq1", ctypes.c_float * 4), ("lsq2", ctypes.c_float * 4),
                ("grid1", ctypes.c_float * 4), ("grid2", ctypes.c_float * 4),
                ("pbit_combo", ctypes.c_int), ("rgb_error", ctypes.c_double),
                ("alpha_error", ctypes.c_double), ("alpha_path", ctypes.c_int),
                ("alpha_centroid", ctypes.c_float * 8), ("alpha_count", ctypes.c_uint32 * 8),
                ("alpha_bucket", ctypes.c_uint8 * 16), ("alpha_lsq", ctypes.c_float * 2),
                ("alpha_q", ctypes.c_uint8 * 2)]


FIT_SAME_POINT, FIT_ONE_BUCKET, FIT_LEAST_SQUARES = 1, 2, 3


# ora_bc7_log_stats / fh_bc7_block_stats
LOG_STATS_DTYPE = np.dtype([("path", "<i4"), ("mode", "<i4"), ("num_two_shape", "<u4"),
                            ("num_three_shape", "<u4"), ("mode_estimate", "<f8", (8,)),
                            ("mode_error", "<f8", (8,)), ("two_shape", "<f8", (64,)),
                            ("three_shape", "<f8", (64,))])
STAT_NAMES = ["Zero", "One", "Two", "Three", "Four", "Five", "Six", "Seven"]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.ora_bc7_compress_block.argtypes = [vp, vp, ctypes.POINTER(Settings), ctypes.c_uint32,
                                             ctypes.c_uint64, ctypes.POINTER(BlockStats)]
        L.ora_bc7_compress_job.argtypes = [vp, vp] + [ctypes.c_uint32] * 6 + [
            ctypes.POINTER(Settings), ctypes.POINTER(ctypes.c_uint32)]
        L.ora_bc7_compress_job_logged.argtypes = [vp, vp] + [ctypes.c_uint32] * 6 + [
            ctypes.POINTER(Settings), ctypes.POINTER(ctypes.c_uint32), vp]
        L.ora_bc7_compress_job_selected.argtypes = [vp, vp] + [ctypes.c_uint32] * 6 + [
            ctypes.POINTER(Settings), ctypes.POINTER(ctypes.c_uint32), vp]
        L.ora_bc7_decompress.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint32]
        L.ora_bc7_decompress_spec.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint32]
        L.ora_bc7_decode_block_spec.argtypes = [vp, vp]
        L.ora_psnr.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint32]
        L.ora_psnr.restype = ctypes.c_double
        L.ora_premul_sse.argtypes = [vp, vp, ctypes.c_uint32]
        L.ora_premul_sse.restype = ctypes.c_double
        L.ora_block_premul_err.argtypes = [vp, vp]
        L.ora_block_premul_err.restype = ctypes.c_uint64
        L.ora_quantize_channel.argtypes = [ctypes.c_uint8, ctypes.c_uint8, ctypes.c_int]
        L.ora_quantize_channel.restype = ctypes.c_uint8
        L.ora_to_pixel.argtypes = [vp, ctypes.c_uint32, ctypes.c_int]
        L.ora_to_pixel.restype = ctypes.c_uint32
        L.ora_length.argtypes = [vp]
        L.ora_length.restype = ctypes.c_float
        L.ora_power_method.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_float)]
        L.ora_power_method.restype = ctypes.c_int
        L.ora_quantized_error.argtypes = [vp, vp, ctypes.c_int, vp, vp, ctypes.c_int,
                                          ctypes.c_uint32, vp, vp]
        L.ora_quantized_error.restype = ctypes.c_double
        L.ora_bc7_box_selection.argtypes = [vp] + [ctypes.POINTER(ctypes.c_uint32)] * 4
        L.ora_bc7_box_selection_m.argtypes = [vp, ctypes.c_uint32] + \
            [ctypes.POINTER(ctypes.c_uint32)] * 4
        L.ora_table_mode5.restype = ctypes.POINTER(ctypes.c_uint8)
        L.ora_table_dxt6.restype = ctypes.POINTER(ctypes.c_uint8)
        L.ora_single_color.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, vp,
                                       ctypes.POINTER(ctypes.c_int)]
        L.ora_single_color.restype = ctypes.c_double
        L.ora_table_partitions.argtypes = [vp] * 5
        L.ora_bc7_exhaustive_label.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
        L.ora_bc7_exhaustive_label.restype = ctypes.c_int
        L.ora_bc7_encode_candidate.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
        L.ora_bc7_encode_candidate.restype = ctypes.c_double
        L.ora_bc7_encode_learned.argtypes = [vp, vp, ctypes.c_int, ctypes.c_uint32,
                                             ctypes.c_uint64, ctypes.c_uint64, vp]
        L.ora_counter_rng.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.ora_counter_rng.restype = ctypes.c_uint32
        L.ora_fit_key.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32]
        L.ora_fit_key.restype = ctypes.c_uint32
        L.ora_dxt_decompress.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, vp]
        L.ora_etc1_decompress.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, vp]
        L.ora_block_errors.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint32, vp]
        L.ora_block_errors.restype = ctypes.c_uint64
        L.ora_principal_axis.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
        L.ora_clamp_to_grid.argtypes = [ctypes.c_int, vp, vp, ctypes.POINTER(ctypes.c_int)]
        L.ora_fit_rgb.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(FitTrace), vp, vp, vp,
                                  ctypes.POINTER(ctypes.c_int)]
        L.ora_fit_rgb.restype = ctypes.c_double
        L.ora_fit_alpha.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(FitTrace), vp, vp, vp, vp]
        L.ora_fit_alpha.restype = ctypes.c_double
        L.ora_bc7_pack.argtypes = [ctypes.c_int] * 4 + [vp] * 5
        L.ora_quantized_error_w.argtypes = [vp, vp, ctypes.c_int, vp, vp, ctypes.c_int,
                                            ctypes.c_uint32, vp, vp, vp]
        L.ora_quantized_error_w.restype = ctypes.c_double
        L.ora_single_color_m.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, vp, vp,
                                         vp, ctypes.POINTER(ctypes.c_int)]
        L.ora_single_color_m.restype = ctypes.c_double
        L.ora_fit_rgb_m.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_uint32, ctypes.POINTER(FitTrace), vp, vp, vp,
                                    ctypes.POINTER(ctypes.c_int)]
        L.ora_fit_rgb_m.restype = ctypes.c_double
        L.ora_fit_alpha_m.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_uint32, ctypes.POINTER(FitTrace), vp, vp, vp,
                                      vp]
        L.ora_fit_alpha_m.restype = ctypes.c_double
        L.ora_anneal_point_count.argtypes = [ctypes.c_int]
        L.ora_anneal_point_count.restype = ctypes.c_uint64
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def bc7_compress(img, quality=0, seed=0, rng=RNG_COUNTER, lcg_seed=0, block_modes=0xFF,
                 rect=None, wm_counter=0xFFFFFFFF, out=None, q0_guard=False, metric=0):
    """BPTCC::Compress over an image or sub-rectangle.  Returns (bytes, wm).
    q0_guard adds the quality guard of the GPU encoder (not in the reference);
    metric is m_ErrorMetric (0 uniform, 1 nonuniform)."""
    img = np.ascontiguousarray(img, dtype=np.uint32)
    h, w = img.shape
    if out is None:
        out = np.zeros((h // 4) * (w // 4) * 16, dtype=np.uint8)
    lcg = ctypes.c_uint32(lcg_seed)
    s = Settings(quality, block_modes, rng, ctypes.pointer(lcg), seed, int(q0_guard), metric)
    wm = ctypes.c_uint32(wm_counter)
    x0, y0, x1, y1 = rect if rect is not None else (0, 0, w, h)
    lib().ora_bc7_compress_job(_p(img), _p(out), w, h, x0, y0, x1, y1, ctypes.byref(s),
                               ctypes.byref(wm))
    return out, wm.value


def bc7_compress_logged(img, quality=0, seed=0, rng=RNG_COUNTER, lcg_seed=0, block_modes=0xFF,
                        rect=None, wm_counter=0xFFFFFFFF, q0_guard=False, metric=0):
    """BPTCC::CompressWithStats over an image or sub-rectangle.  Returns
    (bytes, wm, stats), stats a LOG_STATS_DTYPE array, one record per block
    processed."""
    img = np.ascontiguousarray(img, dtype=np.uint32)
    h, w = img.shape
    out = np.zeros((h // 4) * (w // 4) * 16, dtype=np.uint8)
    stats = np.zeros((h // 4) * (w // 4), dtype=LOG_STATS_DTYPE)
    lcg = ctypes.c_uint32(lcg_seed)
    s = Settings(quality, block_modes, rng, ctypes.pointer(lcg), seed, int(q0_guard), metric)
    wm = ctypes.c_uint32(wm_counter)
    x0, y0, x1, y1 = rect if rect is not None else (0, 0, w, h)
    lib().ora_bc7_compress_job_logged(_p(img), _p(out), w, h, x0, y0, x1, y1, ctypes.byref(s),
                                      ctypes.byref(wm), _p(stats))
    return out, wm.value, stats


# ora_shape_selection / fh_shape_selection
SHAPE_SELECTION_DTYPE = np.dtype([("num_shapes", "<u4"),
                                  ("shapes", [("num_partitions", "<u4"), ("index", "<u4")],
                                   (10,)),
                                  ("selected_modes", "<u4")])


def bc7_compress_selected(img, selections, quality=0, seed=0, rng=RNG_COUNTER, lcg_seed=0,
                          block_modes=0xFF, rect=None, wm_counter=0xFFFFFFFF, q0_guard=False,
                          metric=0):
    """BPTCC::Compress with a caller's shape selection per processed block.
    Returns (bytes, wm)."""
    img = np.ascontiguousarray(img, dtype=np.uint32)
    sels = np.ascontiguousarray(selections, dtype=SHAPE_SELECTION_DTYPE)
    h, w = img.shape
    out = np.zeros((h // 4) * (w // 4) * 16, dtype=np.uint8)
    lcg = ctypes.c_uint32(lcg_seed)
    s = Settings(quality, block_modes, rng, ctypes.pointer(lcg), seed, int(q0_guard), metric)
    wm = ctypes.c_uint32(wm_counter)
    x0, y0, x1, y1 = rect if rect is not None else (0, 0, w, h)
    lib().ora_bc7_compress_job_selected(_p(img), _p(out), w, h, x0, y0, x1, y1, ctypes.byref(s),
                                        ctypes.byref(wm), _p(sels))
    return out, wm.value


def format_block_stats(rec, block_index):
    """The lines BlockLogger writes for one block (Compressor.cpp:1373-1390,
    1945-1950, 1954-2183): "<index>: BlockStat_<name> -- <value>", values
    written by an ostream with its default format (printf %g for doubles)."""
    def line(name, v):
        v = ("%d" % v) if isinstance(v, (int, np.integer)) else ("%g" % v)
        return "%d: BlockStat_%s -- %s\n" % (block_index, name, v)
    out = []
    for k in range(int(rec["num_two_shape"])):
        out.append(line("TwoShapeEstimate", float(rec["two_shape"][k])))
    for k in range(int(rec["num_three_shape"])):
        out.append(line("ThreeShapeEstimate", float(rec["three_shape"][k])))
    out.append(line("Path", int(rec["path"])))
    out.append(line("Mode", int(rec["mode"])))
    for m in range(8):
        out.append(line("Mode%sEstimate" % STAT_NAMES[m], float(rec["mode_estimate"][m])))
        out.append(line("Mode%sError" % STAT_NAMES[m], float(rec["mode_error"][m])))
    return "".join(out)


def bc7_compress_block(px16, quality=0, seed=0, rng=RNG_COUNTER, lcg_seed=0, wm_index=0,
                       block_id=0, block_modes=0xFF, q0_guard=False, metric=0):
    px = np.ascontiguousarray(px16, dtype=np.uint32)
    out = np.zeros(16, dtype=np.uint8)
    lcg = ctypes.c_uint32(lcg_seed)
    s = Settings(quality, block_modes, rng, ctypes.pointer(lcg), seed, int(q0_guard), metric)
    st = BlockStats()
    lib().ora_bc7_compress_block(_p(px), _p(out), ctypes.byref(s), wm_index, block_id,
                                 ctypes.byref(st))
    return out, st


def bc7_decompress(data, width, height):
    data = np.ascontiguousarray(data, dtype=np.uint8)
    out = np.zeros((height, width), dtype=np.uint32)
    lib().ora_bc7_decompress(_p(data), _p(out), width, height)
    return out


def bc7_decompress_spec(data, width, height):
    """The image decoded by the BC7 specification (mode 4 index mode 1 with
    swapped index widths, unlike the reference decoder)."""
    data = np.ascontiguousarray(data, dtype=np.uint8)
    out = np.zeros((height, width), dtype=np.uint32)
    lib().ora_bc7_decompress_spec(_p(data), _p(out), width, height)
    return out


_ref = None


def ref_lib():
    """oracle/_ref/libfastc_ref.so: the reference's own stb_dxt and rg_etc1
    built from the FasTC tree (None when it has not been built)."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB_PATH):
        L = ctypes.CDLL(REF_LIB_PATH)
        vp = ctypes.c_void_p
        L.ref_dxt_compress.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, vp]
        L.ref_etc1_compress.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, vp]
        _ref = L
    return _ref


def dxt_decompress(data, width, height, dxt5):
    """DXTC::DecompressDXT1 / DecompressDXT5 to a (h, w) uint32 image."""
    data = np.ascontiguousarray(data, dtype=np.uint8)
    out = np.zeros((height, width), dtype=np.uint32)
    lib().ora_dxt_decompress(_p(data), width, height, int(dxt5), _p(out))
    return out


def etc1_decompress(data, width, height):
    """ETCC::Decompress to a (h, w) uint32 image."""
    data = np.ascontiguousarray(data, dtype=np.uint8)
    out = np.zeros((height, width), dtype=np.uint32)
    lib().ora_etc1_decompress(_p(data), width, height, _p(out))
    return out


def block_errors(a, b):
    """Per-block premultiplied errors (raster block order) and their total."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    h, w = a.shape
    out = np.zeros(((h + 3) // 4) * ((w + 3) // 4), dtype=np.uint64)
    total = lib().ora_block_errors(_p(a), _p(b), w, h, _p(out))
    return out, total


def bc7_decode_block_spec(blk):
    blk = np.ascontiguousarray(blk, dtype=np.uint8)
    out = np.zeros(16, dtype=np.uint32)
    lib().ora_bc7_decode_block_spec(_p(blk), _p(out))
    return out


def psnr(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    return lib().ora_psnr(_p(a), _p(b), a.shape[1], a.shape[0])


def bc7_exhaustive_label(px16):
    """(label = mode * 64 + slot, error) of the exhaustive quality-0 search."""
    px = np.ascontiguousarray(px16, dtype=np.uint32)
    e = ctypes.c_uint64(0)
    lab = lib().ora_bc7_exhaustive_label(_p(px), ctypes.byref(e))
    return lab, e.value


def bc7_encode_candidate(px16, mode, slot):
    """(16-byte block, internal error) of one quality-0 candidate."""
    px = np.ascontiguousarray(px16, dtype=np.uint32)
    out = np.zeros(16, dtype=np.uint8)
    e = lib().ora_bc7_encode_candidate(_p(px), mode, slot, _p(out))
    return out, e


def bc7_encode_learned(px16, cands, quality=0, seed=0, block_id=0):
    """The learned search's block for candidates [(mode, slot, entry), ...]."""
    px = np.ascontiguousarray(px16, dtype=np.uint32)
    c = np.array([m | (s << 3) | (e << 9) for m, s, e in cands], dtype=np.int32)
    out = np.zeros(16, dtype=np.uint8)
    lib().ora_bc7_encode_learned(_p(px), _p(c), len(c), quality, seed, block_id, _p(out))
    return out


def block_premul_err(src16, dec16):
    a = np.ascontiguousarray(src16, dtype=np.uint32)
    b = np.ascontiguousarray(dec16, dtype=np.uint32)
    return lib().ora_block_premul_err(_p(a), _p(b))


def premul_sse(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    return lib().ora_premul_sse(_p(a), _p(b), a.size)


# --- candidate-ranking MLP oracle (mlp_oracle.c) ----------------------------

MLP_IN, MLP_H, MLP_OUT = 64, 256, 512
MLP_NPARAMS = 64 * 256 + 256 + 256 * 256 + 256 + 256 * 512 + 512


def _mlp_lib():
    L = lib()
    if not getattr(L, "_mlp_ready", False):
        vp = ctypes.c_void_p
        L.ora_bf16_round.argtypes = [ctypes.c_float]
        L.ora_bf16_round.restype = ctypes.c_float
        L.ora_mlp_valid.argtypes = [ctypes.c_int]
        L.ora_mlp_features.argtypes = [vp, ctypes.c_int, vp]
        L.ora_mlp_forward.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, vp, vp, vp]
        L.ora_mlp_xent.argtypes = [vp, vp, ctypes.c_int, vp]
        L.ora_mlp_xent.restype = ctypes.c_double
        L.ora_mlp_loss_grad.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_int, vp]
        L.ora_mlp_loss_grad.restype = ctypes.c_double
        L.ora_adam_step.argtypes = [vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_float, ctypes.c_float, ctypes.c_float]
        L._mlp_ready = True
    return L


def mlp_valid_mask():
    L = _mlp_lib()
    return np.array([L.ora_mlp_valid(j) for j in range(MLP_OUT)], dtype=bool)


def mlp_features(blocks):
    """(n, 16) uint32 blocks -> (n, 64) float32 inputs (RGBA / 255, bf16)."""
    b = np.ascontiguousarray(blocks, dtype=np.uint32)
    x = np.zeros((b.shape[0], MLP_IN), dtype=np.float32)
    _mlp_lib().ora_mlp_features(_p(b), b.shape[0], _p(x))
    return x


def mlp_forward(params, x, round_bf16=True):
    params = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    B = x.shape[0]
    h1 = np.zeros((B, MLP_H), np.float32)
    h2 = np.zeros((B, MLP_H), np.float32)
    lg = np.zeros((B, MLP_OUT), np.float32)
    _mlp_lib().ora_mlp_forward(_p(params), _p(x), B, int(round_bf16), _p(h1), _p(h2), _p(lg))
    return h1, h2, lg


def mlp_xent(logits, labels, want_grad=True):
    logits = np.ascontiguousarray(logits, dtype=np.float32)
    labels = np.ascontiguousarray(labels, dtype=np.int32)
    B = logits.shape[0]
    d = np.zeros_like(logits) if want_grad else None
    loss = _mlp_lib().ora_mlp_xent(_p(logits), _p(labels), B, _p(d) if want_grad else None)
    return loss, d


def mlp_loss_grad(params, x, labels, round_bf16=True, want_grad=True):
    params = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    labels = np.ascontiguousarray(labels, dtype=np.int32)
    g = np.zeros(MLP_NPARAMS, np.float32) if want_grad else None
    loss = _mlp_lib().ora_mlp_loss_grad(_p(params), _p(x), _p(labels), x.shape[0],
                                        int(round_bf16), _p(g) if want_grad else None)
    return loss, g


def adam_step(w, m, v, g, t, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """In place on float32 arrays w, m, v."""
    g = np.ascontiguousarray(g, dtype=np.float32)
    _mlp_lib().ora_adam_step(_p(w), _p(m), _p(v), _p(g), w.size, t, lr, beta1, beta2, eps)


def principal_axis(px, nparts=1, shape=0, subset=0):
    px = np.ascontiguousarray(px, dtype=np.uint32)
    axis = np.zeros(4, np.float32)
    lib().ora_principal_axis(_p(px), nparts, shape, subset, _p(axis))
    return axis


def clamp_to_grid(mode, p1, p2):
    """ClampEndpointsToGrid: returns (p1, p2, p-bit combo index)."""
    a = np.array(p1, dtype=np.float32)
    b = np.array(p2, dtype=np.float32)
    pbc = ctypes.c_int(0)
    lib().ora_clamp_to_grid(mode, _p(a), _p(b), ctypes.byref(pbc))
    return a, b, pbc.value


def fit_rgb(px, mode, shape=0, subset=0, idx_mode=0, metric=0):
    """One RGB cluster fit at quality 0: (error, p1, p2, indices, combo, trace)."""
    px = np.ascontiguousarray(px, dtype=np.uint32)
    tr = FitTrace()
    p1 = np.zeros(4, np.float32)
    p2 = np.zeros(4, np.float32)
    idx = np.zeros(16, np.uint8)
    pbc = ctypes.c_int(0)
    e = lib().ora_fit_rgb_m(_p(px), mode, shape, subset, idx_mode, metric, ctypes.byref(tr),
                            _p(p1), _p(p2), _p(idx), ctypes.byref(pbc))
    return e, p1, p2, idx, pbc.value, tr


def fit_alpha(px, mode, rot, idx_mode=0, metric=0):
    """One rotation-mode fit at quality 0: (error, p1, p2, indices, alpha
    indices, trace of the rotated RGB fit)."""
    px = np.ascontiguousarray(px, dtype=np.uint32)
    tr = FitTrace()
    p1 = np.zeros(4, np.float32)
    p2 = np.zeros(4, np.float32)
    idx = np.zeros(16, np.uint8)
    aidx = np.zeros(16, np.uint8)
    e = lib().ora_fit_alpha_m(_p(px), mode, rot, idx_mode, metric, ctypes.byref(tr), _p(p1),
                              _p(p2), _p(idx), _p(aidx))
    return e, p1, p2, idx, aidx, tr


def quantized_error(px, p1, p2, nbuckets, mask=0xFFFFFFFF, weights=(1.0, 1.0, 1.0, 1.0),
                    sel=None, pbits=None):
    """RGBACluster::QuantizedError over the pixels sel (all 16 by default)
    with per-channel weights: (error, indices in cluster order)."""
    px = np.ascontiguousarray(px, dtype=np.uint32)
    f1 = np.ascontiguousarray(p1, dtype=np.float32)
    f2 = np.ascontiguousarray(p2, dtype=np.float32)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    idx = np.zeros(16, np.uint8)
    n = 16
    sp = None
    if sel is not None:
        sa = np.ascontiguousarray(sel, dtype=np.uint8)
        n = len(sa)
        sp = _p(sa)
    pb = _p(np.ascontiguousarray(pbits, dtype=np.int32)) if pbits is not None else None
    e = lib().ora_quantized_error_w(_p(px), sp, n, _p(f1), _p(f2), nbuckets, mask, pb, _p(w),
                                    _p(idx))
    return e, idx[:n].copy()


def metric_weights(metric):
    """kErrorMetrics (Compressor.cpp:205-208) as float32."""
    if metric == 1:
        return np.sqrt(np.array([0.3, 0.56, 0.11], np.float32)).tolist() + [1.0]
    return [1.0, 1.0, 1.0, 1.0]


def bc7_pack(mode, shape, rot, idx_mode, ep, pbc, idx, aidx):
    """CompressionMode::Pack of explicit parameters (ep: [3][2][4] floats)."""
    ep = np.ascontiguousarray(ep, dtype=np.float32).reshape(3, 2, 4)
    pbc = np.ascontiguousarray(pbc, dtype=np.uint8).reshape(3)
    idx = np.ascontiguousarray(idx, dtype=np.uint8).reshape(16)
    aidx = np.ascontiguousarray(aidx, dtype=np.uint8).reshape(16)
    out = np.zeros(16, np.uint8)
    lib().ora_bc7_pack(mode, shape, rot, idx_mode, _p(ep), _p(pbc), _p(idx), _p(aidx), _p(out))
    return out
