"""Pins of the CPU oracle's DXT1/DXT5/ETC1 decoders and per-block premultiplied
errors (oracle/decode_oracle.c): hand-computed known answers, the reference's
own rg_etc1 decoder (oracle/_ref, built from the FasTC tree) and its decoded
fixtures, and a float64 restatement of Image::ComputePSNR's per-pixel term."""
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as ora
from tests import ref_lib

HERE = os.path.dirname(os.path.abspath(__file__))


def rgba(r, g, b, a=255):
    return r | (g << 8) | (b << 16) | (a << 24)


def dxt1_block(c0, c1, indices):
    mod = 0
    for i, v in enumerate(indices):
        mod |= v << (2 * i)
    return np.frombuffer(np.array([c0 | (c1 << 16), mod], "<u4").tobytes(), np.uint8)


def dxt5_alpha(a0, a1, indices):
    bits = 0
    for i, v in enumerate(indices):
        bits |= v << (3 * i)
    return np.frombuffer(np.uint64(a0 | (a1 << 8) | (bits << 16)).tobytes(), np.uint8)


def test_dxt1_four_colour_known_answer():
    # red 0xF800 > blue 0x001F: 2/3, 1/3 interpolants truncated
    blk = dxt1_block(0xF800, 0x001F, [i % 4 for i in range(16)])
    got = ora.dxt_decompress(blk, 4, 4, False).ravel()
    want = [rgba(255, 0, 0), rgba(0, 0, 255), rgba(170, 0, 85), rgba(85, 0, 170)]
    assert list(got) == [want[i % 4] for i in range(16)]


def test_dxt1_green_six_bits():
    # green 63 -> 255, green 1 -> 4; (2*255+4)/3 = 171, (255+8)/3 = 87
    blk = dxt1_block(0x07E0, 0x0020, [i % 4 for i in range(16)])
    got = ora.dxt_decompress(blk, 4, 4, False).ravel()
    want = [rgba(0, 255, 0), rgba(0, 4, 0), rgba(0, 171, 0), rgba(0, 87, 0)]
    assert list(got) == [want[i % 4] for i in range(16)]


def test_dxt1_three_colour_mode_keeps_alpha():
    # c0 <= c1: midpoint, and index 3 decodes to black with alpha 255
    for c0, c1 in [(0x001F, 0xF800), (0x1234, 0x1234)]:
        blk = dxt1_block(c0, c1, [i % 4 for i in range(16)])
        got = ora.dxt_decompress(blk, 4, 4, False).ravel()
        if c0 == c1:
            assert got[3] == rgba(0, 0, 0)
            continue
        want = [rgba(0, 0, 255), rgba(255, 0, 0), rgba(127, 0, 127), rgba(0, 0, 0)]
        assert list(got) == [want[i % 4] for i in range(16)]


def test_565_expansion_known_values():
    five = {0: 0, 1: 8, 15: 123, 16: 132, 30: 247, 31: 255}
    six = {0: 0, 1: 4, 31: 125, 32: 130, 63: 255}
    for v, e in five.items():
        got = ora.dxt_decompress(dxt1_block(v << 11 | v, v << 11 | v, [0] * 16), 4, 4, False)
        assert got[0, 0] == rgba(e, 0, e)
    for v, e in six.items():
        got = ora.dxt_decompress(dxt1_block(v << 5, v << 5, [0] * 16), 4, 4, False)
        assert got[0, 0] == rgba(0, e, 0)


def test_dxt5_alpha_modes_known_answer():
    idx = [i % 8 for i in range(16)]
    colour = dxt1_block(0x001F, 0xF800, [2] * 16)
    for a0, a1, pal in [(255, 0, [255, 0, 218, 182, 145, 109, 72, 36]),
                        (0, 255, [0, 255, 51, 102, 153, 204, 0, 255]),
                        (7, 7, [7, 7, 7, 7, 7, 7, 0, 255])]:
        blk = np.concatenate([dxt5_alpha(a0, a1, idx), colour])
        got = ora.dxt_decompress(blk, 4, 4, True).ravel()
        # the colour half of DXT5 always interpolates four colours
        assert all(int(g) & 0xFFFFFF == rgba(85, 0, 170) & 0xFFFFFF for g in got)
        assert [int(g) >> 24 for g in got] == [pal[i] for i in idx]


def test_dxt_edge_blocks_are_clipped():
    rng = np.random.default_rng(1)
    data = rng.integers(0, 256, 3 * 2 * 16, dtype=np.uint8)
    full = ora.dxt_decompress(data, 12, 8, True)
    part = ora.dxt_decompress(data, 10, 7, True)
    assert np.array_equal(part, full[:7, :10])


def test_dxt_solid_roundtrip_through_reference_encoder():
    """stb_dxt (the reference's build) encodes solid colours from its optimal
    5/6-bit match tables; FasTC's decoder brings them back within 2 levels."""
    r = ref_lib.ref()
    v = np.arange(256, dtype=np.uint32)
    img = np.repeat(np.repeat((v | (v << 8) | (v << 16) | (0xFF << 24)).reshape(16, 16),
                              4, 0), 4, 1)
    dec = ora.dxt_decompress(r.dxt(img, False), 64, 64, False)
    d = np.abs(dec.view(np.uint8).astype(int) - img.view(np.uint8).astype(int))
    assert d.max() <= 2


def test_etc1_matches_reference_decoder_on_random_blocks():
    r = ref_lib.ref()
    rng = np.random.default_rng(7)
    data = rng.integers(0, 256, 128 * 128 // 2, dtype=np.uint8)
    want = r.etc1_decode(data, 128, 128)
    assert np.array_equal(ora.etc1_decompress(data, 128, 128), want)
    # clipped edges
    assert np.array_equal(ora.etc1_decompress(data[:8 * 8 * 3], 30, 9),
                          r.etc1_decode(data[:8 * 8 * 3], 30, 9))


def test_etc1_matches_reference_fixture():
    g = np.load(os.path.join(HERE, "golden", "decode.npz"))
    assert np.array_equal(ora.etc1_decompress(g["etc1_random"], 64, 64), g["etc1_random_decoded"])
    assert np.array_equal(ora.etc1_decompress(g["etc1_encoded"], 64, 64), g["etc1_decoded"])


def test_etc1_known_answer_individual_mode():
    # 444 mode, base colours 0x88.. and 0x11.., tables 0 and 7, no flip,
    # all selectors 0 (ETC1 value 0 = +small)
    blk = np.array([0x81, 0x81, 0x81, (0 << 5) | (7 << 2), 0, 0, 0, 0], np.uint8)
    got = ora.etc1_decompress(blk, 4, 4).reshape(4, 4)
    assert got[0, 0] == rgba(0x88 + 2, 0x88 + 2, 0x88 + 2)
    assert got[0, 3] == rgba(0x11 + 47, 0x11 + 47, 0x11 + 47)


def test_block_errors_against_psnr_formula():
    """Per-block errors are Image::ComputePSNR's per-pixel squared differences
    of (a/255)*c, times 255^2, summed over the block's pixels in the image."""
    rng = np.random.default_rng(3)
    h, w = 23, 37
    a = rng.integers(0, 2 ** 32, (h, w), dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 2 ** 32, (h, w), dtype=np.uint64).astype(np.uint32)
    blocks, total = ora.block_errors(a, b)
    pa = a.view(np.uint8).reshape(h, w, 4).astype(np.float64)
    pb = b.view(np.uint8).reshape(h, w, 4).astype(np.float64)
    diff = (pa[..., 3:] / 255.0) * pa[..., :3] - (pb[..., 3:] / 255.0) * pb[..., :3]
    per_px = (diff ** 2).sum(axis=2) * 65025.0
    bw, bh = (w + 3) // 4, (h + 3) // 4
    want = np.zeros(bh * bw)
    for y in range(h):
        for x in range(w):
            want[(y // 4) * bw + x // 4] += per_px[y, x]
    assert np.allclose(blocks.astype(np.float64), want, rtol=1e-12, atol=1e-6)
    assert total == int(blocks.sum())
    mse = (diff ** 2).sum() / (w * h)
    assert math.isclose(ora.psnr(a, b), 10 * math.log10(3 * 255.0 ** 2 / mse), rel_tol=1e-12)
