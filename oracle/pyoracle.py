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
                ("q0_guard", ctypes.c_int)]


class BlockStats(ctypes.Structure):
    _fields_ = [("path", ctypes.c_int), ("mode", ctypes.c_int),
                ("best_error", ctypes.c_double), ("mode_error", ctypes.c_double * 8),
                ("q0_kept", ctypes.c_int)]


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
        L.ora_bc7_decompress.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint32]
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
        L.ora_table_mode5.restype = ctypes.POINTER(ctypes.c_uint8)
        L.ora_table_dxt6.restype = ctypes.POINTER(ctypes.c_uint8)
        L.ora_single_color.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, vp,
                                       ctypes.POINTER(ctypes.c_int)]
        L.ora_single_color.restype = ctypes.c_double
        L.ora_table_partitions.argtypes = [vp] * 5
        L.ora_counter_rng.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.ora_counter_rng.restype = ctypes.c_uint32
        L.ora_fit_key.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32]
        L.ora_fit_key.restype = ctypes.c_uint32
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def bc7_compress(img, quality=0, seed=0, rng=RNG_COUNTER, lcg_seed=0, block_modes=0xFF,
                 rect=None, wm_counter=0xFFFFFFFF, out=None, q0_guard=False):
    """BPTCC::Compress over an image or sub-rectangle.  Returns (bytes, wm).
    q0_guard adds the quality guard of the GPU encoder (not in the reference)."""
    img = np.ascontiguousarray(img, dtype=np.uint32)
    h, w = img.shape
    if out is None:
        out = np.zeros((h // 4) * (w // 4) * 16, dtype=np.uint8)
    lcg = ctypes.c_uint32(lcg_seed)
    s = Settings(quality, block_modes, rng, ctypes.pointer(lcg), seed, int(q0_guard))
    wm = ctypes.c_uint32(wm_counter)
    x0, y0, x1, y1 = rect if rect is not None else (0, 0, w, h)
    lib().ora_bc7_compress_job(_p(img), _p(out), w, h, x0, y0, x1, y1, ctypes.byref(s),
                               ctypes.byref(wm))
    return out, wm.value


def bc7_compress_block(px16, quality=0, seed=0, rng=RNG_COUNTER, lcg_seed=0, wm_index=0,
                       block_id=0, block_modes=0xFF, q0_guard=False):
    px = np.ascontiguousarray(px16, dtype=np.uint32)
    out = np.zeros(16, dtype=np.uint8)
    lcg = ctypes.c_uint32(lcg_seed)
    s = Settings(quality, block_modes, rng, ctypes.pointer(lcg), seed, int(q0_guard))
    st = BlockStats()
    lib().ora_bc7_compress_block(_p(px), _p(out), ctypes.byref(s), wm_index, block_id,
                                 ctypes.byref(st))
    return out, st


def bc7_decompress(data, width, height):
    data = np.ascontiguousarray(data, dtype=np.uint8)
    out = np.zeros((height, width), dtype=np.uint32)
    lib().ora_bc7_decompress(_p(data), _p(out), width, height)
    return out


def bc7_decode_block_spec(blk):
    blk = np.ascontiguousarray(blk, dtype=np.uint8)
    out = np.zeros(16, dtype=np.uint32)
    lib().ora_bc7_decode_block_spec(_p(blk), _p(out))
    return out


def psnr(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    return lib().ora_psnr(_p(a), _p(b), a.shape[1], a.shape[0])


def block_premul_err(src16, dec16):
    a = np.ascontiguousarray(src16, dtype=np.uint32)
    b = np.ascontiguousarray(dec16, dtype=np.uint32)
    return lib().ora_block_premul_err(_p(a), _p(b))


def premul_sse(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    return lib().ora_premul_sse(_p(a), _p(b), a.size)
