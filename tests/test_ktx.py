"""KTX 1.1 writer (fh_ktx_write) against ImageWriterKTX::WriteImage
(IO/src/ImageWriterKTX.cpp:69-160) and the KTX 1.1 layout, on the CPU (the
writer touches no device)."""
import struct

import numpy as np
import pytest

import fastc_amd as fh

IDENT = bytes([0xAB, 0x4B, 0x54, 0x58, 0x20, 0x31, 0x31, 0xBB, 0x0D, 0x0A, 0x1A, 0x0A])


def reference_single_level(internal, base, w, h, data):
    """The bytes ImageWriterKTX writes for one compressed level, spelled out."""
    kv = len(b"KTXorientation\0") + len(b"S=r,T=d\0")   # 23
    total_kv = (kv + 4 + 3) & ~3                           # 28
    head = IDENT + struct.pack("<13I", 0x04030201, 0, 1, 0, internal, base, w, h, 0, 0, 1, 1,
                               total_kv)
    pad = b"KTXorientation"[:total_kv - kv - 4]             # the reference pads with the key
    return head + struct.pack("<I", kv) + b"KTXorientation\0S=r,T=d\0" + pad + \
        struct.pack("<I", len(data)) + data


@pytest.mark.parametrize("fmt,internal,base,bpb", [(3, 0x8E8C, 0x1908, 16),
                                                   (0, 0x83F0, 0x1907, 8),
                                                   (1, 0x83F3, 0x1908, 16)])
def test_single_level_matches_reference_layout(fmt, internal, base, bpb):
    w, h = 16, 8
    data = np.random.default_rng(fmt).integers(0, 256, (w // 4) * (h // 4) * bpb, dtype=np.uint8)
    got = fh.ktx_write(fmt, w, h, [data])
    want = reference_single_level(internal, base, w, h, data.tobytes())
    assert got == want
    # imageSize is the reference's width * height (DXT1: half of it)
    assert struct.unpack_from("<I", got, 64 + 28)[0] == (w * h if bpb == 16 else w * h // 2)


def test_mip_chain_layout():
    w, h = 16, 4
    sizes = []
    lw, lh = w, h
    while True:
        sizes.append(((lw + 3) // 4) * ((lh + 3) // 4) * 16)
        if lw == 1 and lh == 1:
            break
        lw, lh = max(1, lw // 2), max(1, lh // 2)
    assert len(sizes) == 5  # 16x4, 8x2, 4x1, 2x1, 1x1
    levels = [np.full(n, i, np.uint8) for i, n in enumerate(sizes)]
    got = fh.ktx_write(3, w, h, levels)
    assert struct.unpack_from("<I", got, 56)[0] == len(sizes)  # numberOfMipmapLevels
    off = 64 + struct.unpack_from("<I", got, 60)[0]
    for i, n in enumerate(sizes):
        assert struct.unpack_from("<I", got, off)[0] == n
        assert got[off + 4:off + 4 + n] == bytes([i]) * n
        off += 4 + n
    assert off == len(got)


def test_etc1_and_errors():
    data = np.zeros(8, np.uint8)
    got = fh.ktx_write(2, 4, 4, [data])
    assert struct.unpack_from("<II", got, 28) == (0x8D64, 0x1907)  # GL_ETC1_RGB8_OES, GL_RGB
    with pytest.raises(fh.FastHipError):
        fh.ktx_write(9, 4, 4, [data])            # unknown format
    with pytest.raises(fh.FastHipError):
        fh.ktx_write(3, 4, 4, [data, data, data, data])  # 4x4 has 3 levels
