"""C-ABI boundary checks that need no GPU: the library loads, exports every
function include/fasthip.h declares, and reports CompressImageData's
validation failures (Core/src/TexComp.cpp:440-496) as status codes."""
import ctypes
import os
import re

import numpy as np
import pytest

import fastc_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "fasthip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_ ]+\*?\s*(fh_[a-z0-9_]+)\(", text, re.M)))


def test_library_exports_header_symbols():
    L = fastc_amd.lib()
    names = declared_functions()
    assert set(names) >= {"fh_compress", "fh_compress_device", "fh_decompress",
                          "fh_compressed_size", "fh_last_error", "fh_device_available"}
    for n in names:
        assert hasattr(L, n), n
    assert set(fastc_amd.EXPORTED_SYMBOLS) <= set(names)


@pytest.mark.parametrize("fmt,w,h,want", [(0, 256, 256, 256 * 256 // 2),
                                          (1, 256, 256, 256 * 256),
                                          (2, 64, 32, 64 * 32 // 2),
                                          (3, 4096, 4096, 4096 * 4096),
                                          (3, 6, 6, 4 * 16)])
def test_compressed_size(fmt, w, h, want):
    """CompressedImage::GetCompressedSize (Core/src/CompressedImage.cpp:136-149)."""
    assert fastc_amd.compressed_size(fmt, w, h) == want


def _status(img, fmt=3, out_size=None, w=None, h=None):
    L = fastc_amd.lib()
    h0, w0 = img.shape
    w = w0 if w is None else w
    h = h0 if h is None else h
    size = fastc_amd.compressed_size(fmt, w, h) if fmt in (0, 1, 2, 3) else 16
    out = np.zeros(max(size, 16), np.uint8)
    j = fastc_amd._job(fmt, img.ctypes.data, out.ctypes.data,
                       size if out_size is None else out_size, w, h, None, 0, 0, 0xFF, None)
    return L.fh_compress(ctypes.byref(j))


def test_validation_errors_match_reference_checks():
    img = np.zeros((8, 8), np.uint32)
    assert _status(img, w=6) == 3          # not a multiple of the block size
    assert _status(img, fmt=9) == 4        # unknown format
    assert _status(img, out_size=10) == 5  # "Not enough space for compressed data!"
    assert b"space" in fastc_amd.lib().fh_last_error()
    L = fastc_amd.lib()
    j = fastc_amd._job(3, None, img.ctypes.data, 1024, 8, 8, None, 0, 0, 0xFF, None)
    assert L.fh_compress(ctypes.byref(j)) == 2  # "No data sent to compress!"


def test_misaligned_rectangle_rejected():
    img = np.zeros((8, 8), np.uint32)
    out = np.zeros(64, np.uint8)
    j = fastc_amd._job(3, img.ctypes.data, out.ctypes.data, 64, 8, 8, (2, 0, 8, 8), 0, 0,
                       0xFF, None)
    assert fastc_amd.lib().fh_compress(ctypes.byref(j)) == 1


def test_no_device_is_reported_not_emulated():
    """Without a GPU a valid job fails with FH_ERR_NO_DEVICE: there is no CPU
    fallback inside the library."""
    if fastc_amd.device_available():
        pytest.skip("a HIP device is present")
    img = np.zeros((8, 8), np.uint32)
    assert _status(img) == 7
    with pytest.raises(fastc_amd.FastHipError):
        fastc_amd.compress(img)


def test_noalloc_rejects_what_needs_allocation():
    """fh_compress_device_noalloc validates before touching a device: jobs
    over its 16384-block work area, other formats and the learned search are
    refused with a message."""
    L = fastc_amd.lib()
    img = np.zeros((4, 4), np.uint32)
    out = np.zeros(16, np.uint8)
    j = fastc_amd._job(3, img.ctypes.data, out.ctypes.data, 1 << 30, 1024, 1024, None, 0, 0,
                       0xFF, None)
    assert L.fh_compress_device_noalloc(ctypes.byref(j), None) == 1
    assert b"16384" in L.fh_last_error()
    j = fastc_amd._job(0, img.ctypes.data, out.ctypes.data, 16, 4, 4, None, 0, 0, 0xFF, None)
    assert L.fh_compress_device_noalloc(ctypes.byref(j), None) == 6
    j = fastc_amd._job(3, img.ctypes.data, out.ctypes.data, 16, 4, 4, None, 0, 0, 0xFF,
                       ctypes.c_uint32(0))
    assert L.fh_compress_device_noalloc(ctypes.byref(j), None) == 6


def test_modes_that_leave_no_candidate_are_rejected():
    """The reference asserts when a searched block has no mode to try
    (Compressor.cpp:1857): a mode mask without any of modes 0-7, or a caller
    selection that allows nothing on a block that is neither one colour nor
    transparent, is an invalid argument (checked before any device work)."""
    L = fastc_amd.lib()
    img = np.arange(64, dtype=np.uint32).reshape(8, 8) | np.uint32(0xFF000000)
    out = np.zeros(64, np.uint8)
    j = fastc_amd._job(3, img.ctypes.data, out.ctypes.data, 64, 8, 8, None, 0, 0, 0x100, None)
    assert L.fh_compress(ctypes.byref(j)) == 1
    assert b"no BC7 mode" in L.fh_last_error()
    sels = np.zeros(4, dtype=fastc_amd.SHAPE_SELECTION_DTYPE)
    sels["num_shapes"] = 1
    sels["shapes"]["num_partitions"][:, 0] = 2
    sels["selected_modes"] = 1 << 1
    sels["selected_modes"][2] = 1 << 0  # mode 0 needs three subsets: nothing to try
    j = fastc_amd._job(3, img.ctypes.data, out.ctypes.data, 64, 8, 8, None, 0, 0, 0xFF, None)
    assert L.fh_compress_with_selection(ctypes.byref(j), sels.ctypes.data, 4) == 1
    assert b"no mode" in L.fh_last_error()
    # the same empty selection on a one-colour block is never consulted
    img[4:8, 0:4] = 0xFF123456
    if not fastc_amd.device_available():
        assert L.fh_compress_with_selection(ctypes.byref(j), sels.ctypes.data, 4) == 7


def test_error_metric_validation():
    """m_ErrorMetric: 0 and 1 exist (BPTCCompressor.h:73-77); the nonuniform
    metric runs the reference search, so it is refused with the learned
    search and with CompressWithStats, before any device work."""
    L = fastc_amd.lib()
    img = np.zeros((8, 8), np.uint32)
    out = np.zeros(64, np.uint8)
    j = fastc_amd._job(3, img.ctypes.data, out.ctypes.data, 64, 8, 8, None, 0, 0, 0xFF, None,
                       metric=2)
    assert L.fh_compress(ctypes.byref(j)) == 1
    assert b"metric" in L.fh_last_error()
    rk = fastc_amd.FhJob()  # a non-null ranker handle is never dereferenced here
    j = fastc_amd._job(3, img.ctypes.data, out.ctypes.data, 64, 8, 8, None, 0, 0, 0xFF, None,
                       metric=1)
    j.flags |= fastc_amd.FH_FLAG_LEARNED
    j.ranker = ctypes.cast(ctypes.pointer(rk), ctypes.c_void_p)
    assert L.fh_compress(ctypes.byref(j)) == 6
    assert b"uniform" in L.fh_last_error()
    stats = np.zeros(4, dtype=fastc_amd.BLOCK_STATS_DTYPE)
    j = fastc_amd._job(3, img.ctypes.data, out.ctypes.data, 64, 8, 8, None, 0, 0, 0xFF, None,
                       metric=1)
    assert L.fh_compress_with_stats(ctypes.byref(j), stats.ctypes.data, 4) == 6
