"""GPU decoders (DXT1/DXT5/ETC1/BPTC) and the fused encoded-error / PSNR
reductions against the CPU oracle, called through the C ABI.

The oracle's decoders are pinned in tests/test_decode_oracle.py (known
answers; the reference's own rg_etc1 decoder); the BC7 decoder in
tests/test_oracle_bc7.py.  Errors are exact integers, so the comparisons are
exact except PSNR, which the reference sums in float64 pixel by pixel."""
import math
import os

import numpy as np
import pytest
import torch

from fastc_amd import synth

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FORMATS = {"dxt1": (0, 8), "dxt5": (1, 16), "etc1": (2, 8), "bptc": (3, 16)}


def oracle_decode(oracle, fmt, data, w, h):
    if fmt in (0, 1):
        return oracle.dxt_decompress(data, w, h, fmt == 1)
    if fmt == 2:
        return oracle.etc1_decompress(data, w, h)
    return oracle.bc7_decompress(data, w, h)


def random_blocks(fmt, w, h, seed):
    bs = FORMATS[fmt][1]
    n = ((w + 3) // 4) * ((h + 3) // 4)
    return np.random.default_rng(seed).integers(0, 256, n * bs, dtype=np.uint8)


@pytest.mark.parametrize("w,h", [(64, 64), (30, 9), (1024, 512)])
@pytest.mark.parametrize("name", list(FORMATS))
def test_decode_random_blocks(fh, oracle, name, w, h):
    """Arbitrary block bytes (every DXT/ETC1 bit pattern is a valid block;
    random BC7 bytes cover all modes and the reserved one) decode bit-exactly."""
    print(f"decode {name} {w}x{h}", flush=True)
    fmt = FORMATS[name][0]
    data = random_blocks(name, w, h, 17 + fmt)
    got = fh.decompress(data, w, h, fmt=fmt)
    want = oracle_decode(oracle, fmt, data, w, h)
    bad = np.argwhere(got != want)
    assert bad.size == 0, f"{len(bad)} pixels differ, first at {bad[:4].tolist()}"


def test_etc1_decode_reference_fixture(fh):
    """The reference's rg_etc1 decoder output, committed."""
    g = np.load(os.path.join(HERE, "golden", "decode.npz"))
    assert np.array_equal(fh.decompress(g["etc1_random"], 64, 64, fmt=2), g["etc1_random_decoded"])
    assert np.array_equal(fh.decompress(g["etc1_encoded"], 64, 64, fmt=2), g["etc1_decoded"])


@pytest.mark.parametrize("name", list(FORMATS))
def test_decode_device_unaligned_and_odd_width(fh, oracle, name):
    """Device entry point on an output that is not 16-byte aligned and a width
    that is not a multiple of 4 (the scalar store path)."""
    fmt = FORMATS[name][0]
    w, h = 37, 21
    data = random_blocks(name, w, h, 5)
    want = oracle_decode(oracle, fmt, data, w, h)
    d_in = torch.from_numpy(data).cuda()
    out = torch.zeros(w * h + 1, dtype=torch.int32, device="cuda")
    fh.decompress_device(d_in.data_ptr(), out.data_ptr() + 4, w, h, fmt=fmt)
    torch.cuda.synchronize()
    got = out[1:].cpu().numpy().view(np.uint32).reshape(h, w)
    assert np.array_equal(got, want)


def _errors(fh, orig, enc, fmt, w, h):
    d_orig = torch.from_numpy(orig.view(np.int32)).cuda()
    d_enc = torch.from_numpy(enc).cuda()
    nb = ((w + 3) // 4) * ((h + 3) // 4)
    blocks = torch.zeros(nb, dtype=torch.int64, device="cuda")
    sse = torch.full((2,), -1, dtype=torch.int64, device="cuda")
    fh.block_error_device(d_orig.data_ptr(), d_enc.data_ptr(), w, h, sse.data_ptr(),
                          blocks.data_ptr(), fmt=fmt)
    torch.cuda.synchronize()
    s = sse.cpu().numpy().view(np.uint64)
    return blocks.cpu().numpy().view(np.uint64), int(s[0]) + (int(s[1]) << 64)


@pytest.mark.parametrize("w,h", [(64, 64), (37, 21)])
@pytest.mark.parametrize("name", list(FORMATS))
def test_block_error_matches_oracle(fh, oracle, name, w, h):
    print(f"block error {name} {w}x{h}", flush=True)
    fmt = FORMATS[name][0]
    orig = synth.generate((w + 3) // 4 * 4, (h + 3) // 4 * 4, 9)[:h, :w].copy()
    rng = np.random.default_rng(2)
    alpha = rng.integers(0, 256, orig.shape, dtype=np.uint32) << 24
    orig = (orig & np.uint32(0x00FFFFFF)) | np.where(rng.random(orig.shape) < 0.5,
                                                     np.uint32(0xFF000000), alpha)
    data = random_blocks(name, w, h, 31 + fmt)
    blocks, total = _errors(fh, orig, data, fmt, w, h)
    want_blocks, want_total = oracle.block_errors(orig, oracle_decode(oracle, fmt, data, w, h))
    assert np.array_equal(blocks, want_blocks)
    assert total == want_total


def test_image_error_and_psnr(fh, oracle):
    w, h = 257, 131
    a = synth.generate(260, 132, 4)[:h, :w].copy()
    rng = np.random.default_rng(8)
    b = (a ^ rng.integers(0, 8, a.shape, dtype=np.uint32) *
         np.uint32(0x01010101)).astype(np.uint32)
    want_blocks, want_total = oracle.block_errors(a, b)
    da = torch.from_numpy(a.view(np.int32)).cuda()
    db = torch.from_numpy(b.view(np.int32)).cuda()
    blocks = torch.zeros(want_blocks.size, dtype=torch.int64, device="cuda")
    sse = torch.zeros(2, dtype=torch.int64, device="cuda")
    fh.image_error_device(da.data_ptr(), db.data_ptr(), w, h, sse.data_ptr(), blocks.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(blocks.cpu().numpy().view(np.uint64), want_blocks)
    s = sse.cpu().numpy().view(np.uint64)
    assert int(s[0]) == want_total and int(s[1]) == 0
    # the reference sums float64 terms pixel by pixel; the integer total
    # agrees to rounding
    ref = oracle.psnr(a, b)
    assert math.isclose(fh.psnr(a, b), ref, rel_tol=1e-12)
    assert math.isclose(fh.psnr_from_sse(s, w, h), ref, rel_tol=1e-12)
    assert fh.psnr(a, a) == math.inf


def test_full_size_bc7_error_consistency(fh, oracle):
    """4096x4096: the fused error of an encoded BC7 surface equals the image
    error of its decoded image, and its PSNR matches the reference formula on
    the decoded image."""
    print("bc7 4096 fused error", flush=True)
    n = 4096
    img = synth.generate(n, n, 3)
    d_img = torch.from_numpy(img.view(np.int32)).cuda()
    enc = torch.zeros(n * n, dtype=torch.uint8, device="cuda")
    fh.compress_device(d_img.data_ptr(), enc.data_ptr(), n * n, n, n, fmt=3, quality=0)
    dec = torch.zeros(n * n, dtype=torch.int32, device="cuda")
    fh.decompress_device(enc.data_ptr(), dec.data_ptr(), n, n, fmt=3)
    nb = (n // 4) ** 2
    b1 = torch.zeros(nb, dtype=torch.int64, device="cuda")
    b2 = torch.zeros(nb, dtype=torch.int64, device="cuda")
    s1 = torch.zeros(2, dtype=torch.int64, device="cuda")
    s2 = torch.zeros(2, dtype=torch.int64, device="cuda")
    fh.block_error_device(d_img.data_ptr(), enc.data_ptr(), n, n, s1.data_ptr(), b1.data_ptr(),
                          fmt=3)
    fh.image_error_device(d_img.data_ptr(), dec.data_ptr(), n, n, s2.data_ptr(), b2.data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(b1, b2) and torch.equal(s1, s2)
    s = s1.cpu().numpy().view(np.uint64)
    assert int(s[0]) == int(b1.cpu().numpy().view(np.uint64).sum(dtype=np.uint64))
    decoded = dec.cpu().numpy().view(np.uint32).reshape(n, n)
    assert math.isclose(fh.psnr_from_sse(s, n, n), oracle.psnr(img, decoded), rel_tol=1e-12)


def test_bc7_spec_decode(fh, oracle):
    """FH_DECODE_SPEC: BPTC decoded by the BC7 specification equals the
    oracle's specification decoder (pinned by an independent reader in
    tests/test_oracle_fit.py) on random blocks of every mode, index-mode-1
    mode-4 blocks included; the fused error of the same decode matches."""
    rng = np.random.default_rng(71)
    n = 4096
    raw = rng.integers(0, 256, (n, 16), dtype=np.uint8)
    for k in range(n):  # one mode per block, round robin; every 4th mode-4 block index mode 1
        mode = k % 8
        v = int.from_bytes(raw[k].tobytes(), "little")
        v = (v >> (mode + 1) << (mode + 1)) | (1 << mode)
        if mode == 4 and k % 3 == 0:
            v |= 1 << 7
        raw[k] = np.frombuffer(v.to_bytes(16, "little"), np.uint8)
    w, h = 512, 128
    data = raw.reshape(-1)
    got = fh.decompress(data, w, h, fmt=3 | fh.DECODE_SPEC)
    want = np.zeros((h, w), np.uint32)
    for k in range(n):
        by, bx = divmod(k, w // 4)
        want[4 * by:4 * by + 4, 4 * bx:4 * bx + 4] = \
            oracle.bc7_decode_block_spec(raw[k]).reshape(4, 4)
    assert np.array_equal(got, want)
    ref = fh.decompress(data, w, h, fmt=3)
    assert not np.array_equal(ref, got)  # the mode-4 index-mode-1 blocks differ
    img = synth.generate(w, h, 5)
    d_img = torch.from_numpy(img.view(np.int32)).cuda()
    d_enc = torch.from_numpy(data.copy()).cuda()
    sse = torch.zeros(2, dtype=torch.int64, device="cuda")
    blk = torch.zeros(n, dtype=torch.int64, device="cuda")
    fh.block_error_device(d_img.data_ptr(), d_enc.data_ptr(), w, h, sse.data_ptr(),
                          blk.data_ptr(), fmt=3 | fh.DECODE_SPEC)
    torch.cuda.synchronize()
    be = blk.cpu().numpy().view(np.uint64)
    for k in range(0, n, 97):
        by, bx = divmod(k, w // 4)
        src = img[4 * by:4 * by + 4, 4 * bx:4 * bx + 4].reshape(16)
        assert be[k] == oracle.block_premul_err(src, want[4 * by:4 * by + 4,
                                                          4 * bx:4 * bx + 4].reshape(16))
