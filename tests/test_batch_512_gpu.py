"""The texture batch at its benchmark workload: 512 textures (seed 4) of
power-of-two sides 64..2048 with full mip chains, 60% BPTC at quality 50,
25% DXT5, 15% ETC1, run as one autotuned graph launch.

Every DXT5 and ETC1 surface is checked against the reference's own stb_dxt
and rg_etc1 encoders (oracle/_ref), bit for bit.  BPTC blocks are sampled
(the first and last blocks of surfaces, and random ones) and checked
against the oracle with the batch's annealing keys (a block's index among
the batch's BPTC blocks) and watermark entries (solid blocks counted across
BPTC surfaces in submission order)."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from fastc_amd import batch_workload as bw
from fastc_amd import synth
from tests import ref_lib

pytestmark = pytest.mark.gpu


def test_batch_512_against_reference_encoders(fh, oracle):
    import torch
    ref = ref_lib.ref()
    tex = bw.make_batch(count=512, seed=4)
    surfs, keep = [], []
    for t in tex:
        for lv in t["levels"]:
            h, w = lv.shape
            d_in = torch.from_numpy(lv.view(np.int32)).cuda()
            d_out = torch.zeros(fh.compressed_size(t["format"], w, h), dtype=torch.uint8,
                                device="cuda")
            keep.append((t["format"], lv, d_in, d_out))
            surfs.append((t["format"], d_in.data_ptr(), d_out.data_ptr(), w, h))
    print(f"batch of {len(tex)} textures, {len(surfs)} surfaces, q=50", flush=True)
    wm = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    batch = fh.Batch(surfs, quality=50, seed=4, wm_counter_ptr=wm.data_ptr())
    tune = batch.autotune(stream=stream)
    wm.fill_(-1)
    batch.run(use_graph=True, stream=stream)
    torch.cuda.synchronize()
    got = [d_out.cpu().numpy() for _, _, _, d_out in keep]
    print(f"batch done (tune {tune})", flush=True)

    def simple(i):
        fmt, lv = keep[i][0], keep[i][1]
        want = ref.dxt(lv, True) if fmt == bw.FMT_DXT5 else ref.etc1(lv)
        return i, np.array_equal(got[i], want)

    idx = [i for i, k in enumerate(keep) if k[0] != bw.FMT_BPTC]
    # stb_dxt fills its tables on first use behind an unsynchronised flag,
    # writing intermediate values as it searches (stb_dxt.h:121-147,
    # 623-629): threads calling it for the first time at once can read a
    # half-built table.  One call before the threads start builds them.
    ref.dxt(np.zeros((4, 4), np.uint32), True)
    with ThreadPoolExecutor(16) as ex:
        bad = [i for i, ok in ex.map(simple, idx) if not ok]
    assert not bad, f"{len(bad)} DXT5/ETC1 surfaces differ, first {bad[:5]}"

    # BPTC: global block numbering and watermark order of the batch
    bptc = [i for i, k in enumerate(keep) if k[0] == bw.FMT_BPTC]
    blocks = np.concatenate([synth.blocks(keep[i][1]) for i in bptc])
    enc = np.concatenate([got[i].reshape(-1, 16) for i in bptc])
    solid = (blocks == blocks[:, :1]).all(axis=1)
    nsolid_before = np.concatenate([[0], np.cumsum(solid)[:-1]])
    assert int(wm.item()) & 0xFFFFFFFF == (int(solid.sum()) - 1) % 9
    starts = np.concatenate([[0], np.cumsum([synth.blocks(keep[i][1]).shape[0]
                                             for i in bptc])])
    rng = np.random.default_rng(4)
    picks = np.unique(np.concatenate([starts[:-1][::7], starts[1:][::7] - 1,
                                      rng.integers(0, blocks.shape[0], 1500)]))

    def bc7(b):
        ref_blk, _ = oracle.bc7_compress_block(blocks[b], quality=50, seed=4, block_id=int(b),
                                               wm_index=int(nsolid_before[b] % 9),
                                               q0_guard=True)
        return int(b), np.array_equal(enc[b], ref_blk)

    with ThreadPoolExecutor(16) as ex:
        bad = [b for b, ok in ex.map(bc7, picks) if not ok]
    print(f"checked {len(idx)} DXT5/ETC1 surfaces and {len(picks)} of {blocks.shape[0]} "
          f"BPTC blocks", flush=True)
    assert not bad, f"{len(bad)} BPTC blocks differ, first {bad[:8]}"
