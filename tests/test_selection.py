"""A caller's shape selection (BPTCC::CompressionSettings::m_ShapeSelectionFn,
BPTCEncoder/src/Compressor.cpp:1840-1858) through fh_compress_with_selection.

CPU: the oracle's selection path is pinned by BoxSelection's own choices
(feeding them back gives Compress's bytes), by duplicate entries (which cannot
win CompressClusters' strict comparison at quality 0) and by an empty
selection (single-subset modes only); the C ABI's argument checks need no GPU.
GPU: random selections of up to five shapes, at quality 0 and annealed, equal
the oracle bit for bit."""
import ctypes

import numpy as np
import pytest

import fastc_amd
from fastc_amd import synth

SUBSETS = [3, 2, 3, 2, 1, 1, 1, 2]
SLOTS = [1, 1, 1, 1, 8, 4, 1, 1]


def count_candidates(sel, block_modes, quality):
    """CompressClusters' loop (Compressor.cpp:1771-1809) counted the way the
    library slots it: rotation modes 8 and 4, single-subset modes once per
    entry only when annealing."""
    modes = int(sel["selected_modes"]) & block_modes
    nsh = min(5, int(sel["num_shapes"]))
    if nsh == 0:
        nsh = 1
        modes &= ~0x8F
    n = 0
    for mode in range(8):
        if not (modes >> mode) & 1:
            continue
        for e in range(nsh):
            parts, idx = int(sel["shapes"][e]["num_partitions"]), int(sel["shapes"][e]["index"])
            if SUBSETS[mode] != 1 and SUBSETS[mode] != parts:
                continue
            if mode == 0 and idx >= 16:
                continue
            if SUBSETS[mode] == 1 and e > 0 and quality == 0:
                continue
            n += SLOTS[mode]
    return n


def box_selections(oracle, img):
    """BoxSelection's result for every block, as ShapeSelection records."""
    L = oracle.lib()
    h, w = img.shape
    sels = np.zeros((h // 4) * (w // 4), fastc_amd.SHAPE_SELECTION_DTYPE)
    ns, s2, s3, modes = (ctypes.c_uint32() for _ in range(4))
    for k in range(sels.size):
        by, bx = divmod(k, w // 4)
        px = np.ascontiguousarray(img[4 * by:4 * by + 4, 4 * bx:4 * bx + 4]).reshape(16)
        L.ora_bc7_box_selection(px.ctypes.data, ctypes.byref(ns), ctypes.byref(s2),
                                ctypes.byref(s3), ctypes.byref(modes))
        sels[k]["num_shapes"] = ns.value
        sels[k]["shapes"][0] = (2, s2.value)
        sels[k]["shapes"][1] = (3, s3.value)
        sels[k]["selected_modes"] = modes.value
    return sels


def random_selections(n, seed, quality, block_modes=0xFF):
    rng = np.random.default_rng(seed)
    sels = np.zeros(n, fastc_amd.SHAPE_SELECTION_DTYPE)
    for k in range(n):
        while True:
            s = np.zeros(1, fastc_amd.SHAPE_SELECTION_DTYPE)[0]
            s["num_shapes"] = rng.integers(0, 7)
            for e in range(10):
                s["shapes"][e] = (rng.choice([1, 2, 3]), rng.integers(0, 64))
            s["selected_modes"] = rng.integers(1, 256)
            # at least one candidate: the reference asserts on a searched
            # block with none (Compressor.cpp:1857) and the host entry point
            # rejects such a selection
            if 1 <= count_candidates(s, block_modes, quality) <= 31:
                break
        sels[k] = s
    return sels


def test_box_selection_fed_back_gives_compress(oracle):
    img = synth.generate(64, 64, 4)
    sels = box_selections(oracle, img)
    for q in (0, 5):
        a, _ = oracle.bc7_compress(img, quality=q, seed=3, q0_guard=True)
        b, _ = oracle.bc7_compress_selected(img, sels, quality=q, seed=3, q0_guard=True)
        assert np.array_equal(a, b), q


def test_duplicate_entries_cannot_win_at_quality_0(oracle):
    img = synth.generate(32, 32, 6)
    sels = box_selections(oracle, img)
    a, _ = oracle.bc7_compress_selected(img, sels)
    dup = sels.copy()
    for k in range(dup.size):
        dup[k]["shapes"][2:5] = dup[k]["shapes"][0:3]
        dup[k]["num_shapes"] = 5
    b, _ = oracle.bc7_compress_selected(img, dup)
    assert np.array_equal(a, b)


def test_empty_selection_uses_single_subset_modes(oracle):
    img = synth.generate(32, 32, 6)
    sels = np.zeros((32 // 4) ** 2, fastc_amd.SHAPE_SELECTION_DTYPE)
    sels["selected_modes"] = 0xFF
    out, _ = oracle.bc7_compress_selected(img, sels)
    modes = [next(m for m in range(8) if (b >> m) & 1) for b in out[::16]]
    assert set(modes) <= {4, 5, 6}


def _job(img, out, q=0):
    h, w = img.shape
    return fastc_amd._job(3, img.ctypes.data, out.ctypes.data, out.nbytes, w, h, None, q, 0, 0xFF,
                          None, True, None, 4, 0.6)


def test_selection_argument_checks():
    """Checked before any device work: too many candidates, a shape index past
    the partition tables, too few entries, other formats."""
    L = fastc_amd.lib()
    img = np.zeros((8, 8), np.uint32)
    out = np.zeros(64, np.uint8)
    j = _job(img, out, q=1)
    sels = np.zeros(4, fastc_amd.SHAPE_SELECTION_DTYPE)
    sels["selected_modes"] = 0xFF
    sels["num_shapes"] = 4
    sels["shapes"][:, :4] = (1, 0)  # modes 4, 5, 6 four times each with annealing: 52
    assert count_candidates(sels[0], 0xFF, 1) == 52
    assert L.fh_compress_with_selection(ctypes.byref(j), sels.ctypes.data, 4) == 6
    sels["num_shapes"] = 2  # 26
    sels["shapes"][:, 1] = (2, 64)
    assert L.fh_compress_with_selection(ctypes.byref(j), sels.ctypes.data, 4) == 1
    assert L.fh_compress_with_selection(ctypes.byref(j), sels.ctypes.data, 3) == 5
    j.format = 0
    assert L.fh_compress_with_selection(ctypes.byref(j), sels.ctypes.data, 4) == 6


@pytest.mark.gpu
@pytest.mark.parametrize("q,seed", [(0, 1), (0, 2), (6, 3)])
def test_random_selections_bit_exact(fh, oracle, q, seed):
    img = synth.generate(64, 64, 5)
    print(f"selection q{q} seed={seed}", flush=True)
    sels = random_selections((64 // 4) ** 2, seed, q)
    ref, _ = oracle.bc7_compress_selected(img, sels, quality=q, seed=seed, q0_guard=True)
    got = fh.compress_with_selection(img, sels, quality=q, seed=seed,
                                     wm_counter=ctypes.c_uint32(0xFFFFFFFF))
    bad = np.nonzero((got.reshape(-1, 16) != ref.reshape(-1, 16)).any(1))[0]
    assert bad.size == 0, f"{bad.size} blocks differ, first {bad[:8]}"


@pytest.mark.gpu
def test_box_selection_through_the_hook_equals_compress(fh, oracle):
    img = synth.generate(128, 64, 2)
    print("selection = BoxSelection", flush=True)
    sels = box_selections(oracle, img)
    for q in (0, 8):
        a = fh.compress(img, quality=q, seed=5, wm_counter=ctypes.c_uint32(0xFFFFFFFF))
        b = fh.compress_with_selection(img, sels, quality=q, seed=5,
                                       wm_counter=ctypes.c_uint32(0xFFFFFFFF))
        assert np.array_equal(a, b), q


@pytest.mark.gpu
def test_selections_multi_call_job(fh, oracle):
    """More than 2^20 blocks: the second call reads its own selections."""
    w, h = 4104, 4096
    img = synth.generate(w, h, 12)
    print("selection multi-call 4104x4096", flush=True)
    n = (w // 4) * (h // 4)
    sels = np.zeros(n, fastc_amd.SHAPE_SELECTION_DTYPE)
    k = np.arange(n)
    sels["num_shapes"] = 1
    sels["shapes"][:, 0]["num_partitions"] = 2
    sels["shapes"][:, 0]["index"] = k % 64
    sels["selected_modes"] = np.where(k % 3 == 0, 0x82, 0xC2)  # modes 1, 7 (and 6)
    got = fh.compress_with_selection(img, sels, wm_counter=ctypes.c_uint32(0xFFFFFFFF))
    bw = w // 4
    blocks = img.reshape(h // 4, 4, bw, 4).transpose(0, 2, 1, 3).reshape(-1, 16)
    rng = np.random.default_rng(5)
    checked = 0
    for b in np.concatenate([rng.integers(1 << 20, n, 20), rng.integers(0, 1 << 20, 6)]):
        px = blocks[b]
        if (px == px[0]).all() or not (px >> 24).any():
            continue
        ref, _ = oracle.bc7_compress_selected(px.reshape(4, 4), sels[b:b + 1])
        assert np.array_equal(got[16 * b:16 * b + 16], ref), b
        checked += 1
    assert checked >= 15


@pytest.mark.gpu
def test_selection_device_entry_point(fh, oracle):
    """fh_compress_with_selection_device on torch buffers equals the host
    entry point."""
    import torch
    img = synth.generate(64, 64, 7)
    print("selection device entry point", flush=True)
    sels = random_selections((64 // 4) ** 2, 9, 3)
    want = fh.compress_with_selection(img, sels, quality=3, seed=1,
                                      wm_counter=ctypes.c_uint32(0xFFFFFFFF))
    d_img = torch.from_numpy(img.view(np.int32)).cuda()
    d_sel = torch.from_numpy(sels.view(np.uint8).copy()).cuda()
    enc = torch.zeros(64 * 64, dtype=torch.uint8, device="cuda")
    fh.compress_with_selection_device(d_img.data_ptr(), enc.data_ptr(), enc.numel(),
                                      d_sel.data_ptr(), sels.size, 64, 64, quality=3, seed=1,
                                      wm_counter=ctypes.c_uint32(0xFFFFFFFF))
    torch.cuda.synchronize()
    assert np.array_equal(enc.cpu().numpy(), want)
