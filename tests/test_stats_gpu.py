"""CompressWithStats on the GPU (fh_compress_with_stats) against the CPU
oracle's statistics block path (BPTCEncoder/src/Compressor.cpp:1954-2183).

Bit-exact bar: the encoded bytes and every field of every block's statistics
record (estimates are doubles computed in the reference's order) equal the
oracle's, at quality 0 and, with the counter annealing stream on both sides,
at quality > 0, where the second shape entry of modes 4, 5 and 6 anneals on
its own stream.  At 1024x1024 the records are checked through properties
that do not need the oracle.  Every call passes its own watermark counter
(FH_FLAG_CALLER_WATERMARK), starting where a fresh process's gWMVal does, as
the oracle's does."""
import ctypes

import numpy as np
import pytest

from fastc_amd import synth

pytestmark = pytest.mark.gpu


def _wm():
    return ctypes.c_uint32(0xFFFFFFFF)


def _near_opaque(img, seed):
    """Some blocks with alpha 250..254 (opaque to Compress's BoxSelection, not
    to the statistics path) and some fully opaque."""
    img = img.copy()
    rng = np.random.default_rng(seed)
    h, w = img.shape
    for by in range(h // 4):
        for bx in range(w // 4):
            r = rng.random()
            blk = img[4 * by:4 * by + 4, 4 * bx:4 * bx + 4]
            if r < 0.15:
                blk[:] = (blk & 0x00FFFFFF) | (rng.integers(250, 255, (4, 4)).astype(np.uint32)
                                               << 24)
            elif r < 0.3:
                blk[:] = blk | np.uint32(0xFF000000)
    return img


def _compare(got, ref, gst, rst):
    a, b = got.reshape(-1, 16), ref.reshape(-1, 16)
    bad = np.nonzero((a != b).any(1))[0]
    assert bad.size == 0, f"{bad.size} blocks differ, first {bad[:8]}"
    assert gst.size == rst.size
    for f in gst.dtype.names:
        diff = np.nonzero((gst[f] != rst[f]).reshape(gst.size, -1).any(1))[0]
        assert diff.size == 0, f"{f}: {diff.size} records differ, first {diff[:8]}"


@pytest.mark.parametrize("w,h,seed", [(64, 64, 1), (256, 256, 1), (96, 32, 4), (4, 4, 2)])
def test_stats_q0_bit_exact(fh, oracle, w, h, seed):
    img = _near_opaque(synth.generate(w, h, seed), seed)
    print(f"stats q0 {w}x{h} seed={seed}", flush=True)
    ref, ref_wm, rst = oracle.bc7_compress_logged(img)
    wm = _wm()
    got, gst = fh.compress_with_stats(img, wm_counter=wm)
    _compare(got, ref, gst, rst)
    assert wm.value == ref_wm


@pytest.mark.parametrize("q,guard", [(8, True), (50, False)])
def test_stats_annealed_bit_exact(fh, oracle, q, guard):
    """Annealing with both shape entries of the single-subset modes."""
    img = _near_opaque(synth.generate(64, 64, 5), 3)
    print(f"stats q{q} guard={guard}", flush=True)
    ref, _, rst = oracle.bc7_compress_logged(img, quality=q, seed=9, q0_guard=guard)
    got, gst = fh.compress_with_stats(img, quality=q, seed=9, q0_guard=guard, wm_counter=_wm())
    _compare(got, ref, gst, rst)


@pytest.mark.parametrize("modes", [0x0F, 0x30, 0x40, 0x82])
def test_stats_block_mode_masks(fh, oracle, modes):
    img = _near_opaque(synth.generate(64, 64, 6), 6)
    print(f"stats modes {modes:#x}", flush=True)
    ref, _, rst = oracle.bc7_compress_logged(img, block_modes=modes)
    got, gst = fh.compress_with_stats(img, block_modes=modes, wm_counter=_wm())
    _compare(got, ref, gst, rst)


def test_stats_sub_job(fh, oracle):
    """A sub-rectangle (CompressionJob's block-raster rules): records in
    processing order, the rest of the array untouched, bytes only inside."""
    img = synth.generate(64, 64, 2)
    rect = (8, 4, 40, 48)
    print("stats sub-job", flush=True)
    ref, _, rst = oracle.bc7_compress_logged(img, rect=rect)
    got, gst = fh.compress_with_stats(img, rect=rect, wm_counter=_wm())
    # rows 4..48: 14 blocks from x = 8, ten full rows of 16, 10 up to x = 40
    assert int(np.count_nonzero((rst["mode"] != 0) | (rst["path"] != 0))) == 14 + 160 + 10
    _compare(got, ref, gst, rst)


def test_stats_text_matches_oracle(fh, oracle):
    img = synth.generate(32, 32, 3)
    got, gst = fh.compress_with_stats(img, wm_counter=_wm())
    _, _, rst = oracle.bc7_compress_logged(img)
    for k in range(gst.size):
        assert fh.stats_format(gst[k], k) == oracle.format_block_stats(rst[k], k)


def test_stats_1024_properties(fh, oracle):
    """1024x1024: every record is consistent with the bytes written and with
    the block it describes; a sample of blocks matches the oracle."""
    n = 1024
    img = _near_opaque(synth.generate(n, n, 3), 1)
    print("stats 1024 properties", flush=True)
    got, st = fh.compress_with_stats(img, wm_counter=_wm())
    blocks = img.reshape(n // 4, 4, n // 4, 4).transpose(0, 2, 1, 3).reshape(-1, 16)
    first = got.reshape(-1, 16)[:, 0].astype(np.int64)
    low_bit = np.where(first == 0, 8, np.log2(first & -first).astype(np.int64))
    assert np.array_equal(st["mode"], low_bit)
    solid = (blocks == blocks[:, :1]).all(1)
    transparent = ~solid & ((blocks >> 24) == 0).all(1)
    assert np.array_equal(st["path"] == 0, solid)
    assert np.array_equal(st["path"] == 1, transparent)
    assert (st["mode_error"] == -1).all()
    opaque = ((blocks >> 24) == 255).all(1)
    searched = st["path"] >= 2
    assert ((st["num_three_shape"] > 0) == (opaque & searched)).all()
    assert (st["num_two_shape"][searched] > 0).all()
    rng = np.random.default_rng(0)
    for k in rng.choice(st.size, 48, replace=False):
        px = blocks[k].reshape(4, 4)
        ref, _, rst = oracle.bc7_compress_logged(px)
        if st["path"][k] == 0:
            continue  # the watermark entry depends on the image's earlier solid blocks
        assert np.array_equal(got[16 * k:16 * k + 16], ref)
        for f in st.dtype.names:
            assert np.array_equal(st[f][k], rst[f][0]), (k, f)


def test_stats_device_entry_point(fh, oracle):
    """fh_compress_with_stats_device on torch buffers equals the oracle, and
    rejects a short record buffer."""
    import torch
    img = _near_opaque(synth.generate(128, 64, 8), 8)
    print("stats device entry point", flush=True)
    ref, _, rst = oracle.bc7_compress_logged(img, quality=4, seed=2, q0_guard=True)
    d_img = torch.from_numpy(img.view(np.int32)).cuda()
    enc = torch.zeros(128 * 64, dtype=torch.uint8, device="cuda")
    nb = 128 * 64 // 16
    rec = torch.full((nb * fh.BLOCK_STATS_DTYPE.itemsize,), 0x7F, dtype=torch.uint8,
                     device="cuda")
    fh.compress_with_stats_device(d_img.data_ptr(), enc.data_ptr(), enc.numel(), rec.data_ptr(),
                                  nb, 128, 64, quality=4, seed=2, wm_counter=_wm())
    torch.cuda.synchronize()
    _compare(enc.cpu().numpy(), ref, rec.cpu().numpy().view(fh.BLOCK_STATS_DTYPE), rst)
    with pytest.raises(fh.FastHipError):
        fh.compress_with_stats_device(d_img.data_ptr(), enc.data_ptr(), enc.numel(),
                                      rec.data_ptr(), nb - 1, 128, 64, wm_counter=_wm())


def test_stats_multi_call_job(fh, oracle):
    """A job of more than 2^20 blocks runs as consecutive calls: the records
    of the second call land at their blocks (sampled against the oracle)."""
    w, h = 4104, 4096  # 1,050,624 blocks
    img = synth.generate(w, h, 11)
    print("stats multi-call 4104x4096", flush=True)
    got, st = fh.compress_with_stats(img, wm_counter=_wm())
    bw = w // 4
    blocks = img.reshape(h // 4, 4, bw, 4).transpose(0, 2, 1, 3).reshape(-1, 16)
    rng = np.random.default_rng(4)
    picks = np.concatenate([rng.integers((1 << 20) - 64, 1 << 20, 6),
                            rng.integers(1 << 20, st.size, 18), rng.integers(0, 1 << 20, 8)])
    checked = 0
    for k in picks:
        if st["path"][k] == 0:
            continue  # a solid block's watermark entry depends on the blocks before it
        ref, _, rst = oracle.bc7_compress_logged(blocks[k].reshape(4, 4))
        assert np.array_equal(got[16 * k:16 * k + 16], ref), k
        for f in st.dtype.names:
            assert np.array_equal(st[f][k], rst[f][0]), (k, f)
        checked += 1
    assert checked >= 20
