"""The nonuniform error metric (m_ErrorMetric = eErrorMetric_Nonuniform,
BPTCCompressor.h:73-77) on the GPU against the oracle, bit for bit: quality
0 and annealed (the shared counter stream), with and without the quality
guard, block-mode masks, a caller's shape selection, and a batch.  Every
error of this metric is a float sum in the reference's order, so any
reordering of an addition shows up as a different choice somewhere in
these images."""
import ctypes

import numpy as np
import pytest

from fastc_amd import synth
from tests.test_selection import random_selections

pytestmark = pytest.mark.gpu


def _wm():
    return ctypes.c_uint32(0xFFFFFFFF)


def _diff(a, b):
    a = a.reshape(-1, 16)
    b = b.reshape(-1, 16)
    return np.nonzero((a != b).any(axis=1))[0]


@pytest.mark.parametrize("w,h,seed", [(64, 64, 1), (128, 128, 7), (96, 32, 4)])
def test_nonuniform_q0_bit_exact(fh, oracle, w, h, seed):
    img = synth.generate(w, h, seed)
    print(f"nonuniform q0 {w}x{h} seed={seed}", flush=True)
    ref, ref_wm = oracle.bc7_compress(img, quality=0, metric=1)
    wm = _wm()
    got = fh.compress(img, quality=0, wm_counter=wm, metric=fh.eErrorMetric_Nonuniform)
    bad = _diff(got, ref)
    assert bad.size == 0, f"{bad.size} blocks differ, first {bad[:8]}"
    assert wm.value == ref_wm
    # the metric changes the encoder's choices
    uni = fh.compress(img, quality=0, wm_counter=_wm())
    assert _diff(got, uni).size > 0


@pytest.mark.parametrize("q,guard", [(8, False), (50, True), (50, False)])
def test_nonuniform_annealed_bit_exact(fh, oracle, q, guard):
    img = synth.generate(64, 64, 3)
    print(f"nonuniform q={q} guard={guard}", flush=True)
    ref, _ = oracle.bc7_compress(img, quality=q, seed=1234, q0_guard=guard, metric=1)
    got = fh.compress(img, quality=q, seed=1234, wm_counter=_wm(), q0_guard=guard, metric=1)
    bad = _diff(got, ref)
    assert bad.size == 0, f"{bad.size} blocks differ, first {bad[:8]}"


@pytest.mark.parametrize("modes", [0x30, 0xCF, 0x41, 0x80])
def test_nonuniform_block_modes(fh, oracle, modes):
    img = synth.generate(64, 64, 6)
    print(f"nonuniform modes={modes:#x}", flush=True)
    ref, _ = oracle.bc7_compress(img, quality=4, seed=77, block_modes=modes, q0_guard=True,
                                 metric=1)
    got = fh.compress(img, quality=4, seed=77, block_modes=modes, wm_counter=_wm(), metric=1)
    bad = _diff(got, ref)
    assert bad.size == 0, f"{bad.size} blocks differ, first {bad[:8]}"


def test_nonuniform_with_shape_selection(fh, oracle):
    img = synth.generate(64, 64, 8)
    n = (64 // 4) * (64 // 4)
    sels = random_selections(n, 5, 6)
    ref, _ = oracle.bc7_compress_selected(img, sels, quality=6, seed=3, q0_guard=True,
                                          metric=1)
    got = fh.compress_with_selection(img, sels, quality=6, seed=3, wm_counter=_wm(), metric=1)
    bad = _diff(got, ref)
    assert bad.size == 0, f"{bad.size} blocks differ, first {bad[:8]}"


def test_nonuniform_batch(fh, oracle):
    """A batch's BPTC blocks under the nonuniform metric: the oracle's
    encoding of the same blocks in submission order (block ids and
    watermark order as the batch numbers them)."""
    import torch
    from tests.test_batch_gpu import _outputs, _setup
    surfs, keep = _setup(fh, 10, 43)
    wm = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    b = fh.Batch(surfs, quality=5, seed=9, wm_counter_ptr=wm.data_ptr(), metric=1)
    b.run()
    got = _outputs(keep)
    bptc = [lv for fmt, lv, _, _ in keep if fmt == 3]
    assert bptc
    blocks = np.concatenate([synth.blocks(lv) for lv in bptc])
    col = np.ascontiguousarray(blocks.reshape(-1, 4, 4).reshape(-1, 4))
    want, _ = oracle.bc7_compress(col, quality=5, seed=9, q0_guard=True, metric=1)
    want = want.reshape(-1, 16)
    at = 0
    for i, (fmt, lv, _, _) in enumerate(keep):
        if fmt != 3:
            continue
        nb = (lv.shape[0] // 4) * (lv.shape[1] // 4)
        assert np.array_equal(got[i].reshape(-1, 16), want[at:at + nb]), i
        at += nb
