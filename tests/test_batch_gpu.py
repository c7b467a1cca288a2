"""Batched encode (fh_batch_*): every surface of a mixed batch with mip
chains equals the single-surface encode; graph replays equal the eager run;
the autotuner's launch shapes do not change the bytes."""
import ctypes

import numpy as np
import pytest

from fastc_amd import batch_workload as bw
from fastc_amd import synth

pytestmark = pytest.mark.gpu


def _setup(fh, count, seed, max_side=64):
    import torch
    tex = bw.make_batch(count=count, seed=seed, min_side=4, max_side=max_side)
    surfs, keep = [], []
    for t in tex:
        for lv in t["levels"]:
            h, w = lv.shape
            d_in = torch.from_numpy(lv.view(np.int32).copy()).cuda()
            n = fh.compressed_size(t["format"], w, h)
            d_out = torch.zeros(n, dtype=torch.uint8, device="cuda")
            keep.append((t["format"], lv, d_in, d_out))
            surfs.append((t["format"], d_in.data_ptr(), d_out.data_ptr(), w, h))
    return surfs, keep


def _outputs(keep):
    import torch
    torch.cuda.synchronize()
    return [d_out.cpu().numpy().copy() for _, _, _, d_out in keep]


@pytest.mark.parametrize("quality", [0, 6])
def test_batch_matches_single_surface_encodes(fh, quality):
    import torch
    surfs, keep = _setup(fh, 14, 41)
    wm = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    print(f"batch of {len(surfs)} surfaces, q={quality}", flush=True)
    b = fh.Batch(surfs, quality=quality, seed=9, wm_counter_ptr=wm.data_ptr())
    b.run()
    got = _outputs(keep)
    # reference: every BPTC block of the batch as one image, one block wide,
    # in submission order (so block ids and watermark order match)
    bptc = [lv for fmt, lv, _, _ in keep if fmt == 3]
    blocks = np.concatenate([synth.blocks(lv) for lv in bptc]) if bptc else None
    whole = None
    if blocks is not None:
        col = blocks.reshape(-1, 4, 4).reshape(-1, 4)  # (4n, 4) image of n blocks
        whole = fh.compress(np.ascontiguousarray(col), quality=quality, seed=9,
                            wm_counter=ctypes.c_uint32(0xFFFFFFFF)).reshape(-1, 16)
    at = 0
    for i, (fmt, lv, _, _) in enumerate(keep):
        if fmt == 3:
            nb = (lv.shape[0] // 4) * (lv.shape[1] // 4)
            assert np.array_equal(got[i].reshape(-1, 16), whole[at:at + nb]), i
            at += nb
        else:
            want = fh.compress(lv, fmt=fmt)
            assert np.array_equal(got[i], want), i
    assert wm.item() != -1 or blocks is None or not (blocks == blocks[:, :1]).all(1).any()


def test_batch_graph_replays_equal_eager(fh):
    import torch
    surfs, keep = _setup(fh, 10, 42)
    wm = torch.zeros((1,), dtype=torch.int32, device="cuda")
    b = fh.Batch(surfs, quality=4, seed=1, wm_counter_ptr=wm.data_ptr())
    wm.fill_(-1)
    b.run()
    eager = _outputs(keep)
    for rep in range(3):
        for _, _, _, d_out in keep:
            d_out.zero_()
        wm.fill_(-1)
        b.run(use_graph=True)
        got = _outputs(keep)
        for i in range(len(keep)):
            assert np.array_equal(got[i], eager[i]), (rep, i)
        # eager work between replays
        wm.fill_(-1)
        b.run()
        torch.cuda.synchronize()


def test_batch_autotune_keeps_bytes(fh):
    import torch
    surfs, keep = _setup(fh, 8, 43, max_side=128)
    wm = torch.zeros((1,), dtype=torch.int32, device="cuda")
    b = fh.Batch(surfs, quality=3, seed=2, wm_counter_ptr=wm.data_ptr())
    wm.fill_(-1)
    b.run()
    before = _outputs(keep)
    tune = b.autotune()
    print(tune, flush=True)
    assert tune["wg_per_cu"] in (8, 16, 32) and tune["anneal_unroll"] in (1, 2, 4)
    assert tune["simple_wg"] in (64, 128, 256, 512) and tune["trials"] >= 7
    wm.fill_(-1)
    b.run(use_graph=True)
    after = _outputs(keep)
    for i in range(len(keep)):
        assert np.array_equal(after[i], before[i]), i
