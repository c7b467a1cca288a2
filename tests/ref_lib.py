"""Loader for oracle/_ref/libfastc_ref.so: the reference's own stb_dxt and
rg_etc1 encoders built from the FasTC tree (oracle/ref/Makefile).  Tests that
need it skip when it has not been built."""
import ctypes
import os

import numpy as np
import pytest

PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                    "oracle", "_ref", "libfastc_ref.so")


class _Ref:
    def __init__(self):
        L = ctypes.CDLL(PATH)
        vp = ctypes.c_void_p
        L.ref_dxt_compress.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, vp]
        L.ref_etc1_compress.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, vp]
        L.ref_etc1_decompress.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, vp]
        self.L = L

    def dxt(self, img, alpha):
        img = np.ascontiguousarray(img, dtype=np.uint32)
        h, w = img.shape
        out = np.zeros((h // 4) * (w // 4) * (16 if alpha else 8), np.uint8)
        self.L.ref_dxt_compress(img.ctypes.data, w, h, int(alpha), out.ctypes.data)
        return out

    def etc1(self, img):
        img = np.ascontiguousarray(img, dtype=np.uint32)
        h, w = img.shape
        out = np.zeros((h // 4) * (w // 4) * 8, np.uint8)
        self.L.ref_etc1_compress(img.ctypes.data, w, h, out.ctypes.data)
        return out

    def etc1_decode(self, data, w, h):
        out = np.zeros((h, w), np.uint32)
        self.L.ref_etc1_decompress(np.ascontiguousarray(data).ctypes.data, w, h, out.ctypes.data)
        return out


_INST = None


def ref():
    global _INST
    if not os.path.exists(PATH):
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    if _INST is None:
        _INST = _Ref()
    return _INST
