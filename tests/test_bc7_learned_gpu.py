"""Learned candidate pruning (FH_FLAG_LEARNED) against the oracle, bit for
bit: the test reads the ranker's logits from the GPU and reproduces the
pruning decisions.  A confident block (top softmax probability at least the
threshold) is encoded by the oracle from its top-k candidates with the
learned search's decoded-error selection; any other block must be exactly
the full search's block (BoxSelection, every candidate annealed, the
reference's choice, the quality guard)."""
import ctypes

import numpy as np
import pytest

from fastc_amd import synth

pytestmark = pytest.mark.gpu


def _box_candidates(oracle, px, block_modes, quality):
    """The candidate list BoxSelection gives one block, in CompressClusters'
    order: (mode, slot, shape entry)."""
    L = oracle.lib()
    ns, s2, s3, modes = (ctypes.c_uint32() for _ in range(4))
    L.ora_bc7_box_selection(np.ascontiguousarray(px).ctypes.data, ctypes.byref(ns),
                            ctypes.byref(s2), ctypes.byref(s3), ctypes.byref(modes))
    nsh, s2, s3, m_all = ns.value, s2.value, s3.value, modes.value
    m = m_all & block_modes
    out = []
    if not (m_all >> 4) & 1:
        if nsh >= 2:
            if m & 1 and s3 < 16:
                out.append((0, s3, 1))
            if (m >> 2) & 1:
                out.append((2, s3, 1))
        for mode in (1, 3, 7):
            if (m >> mode) & 1:
                out.append((mode, s2, 0))
        if (m >> 6) & 1:
            out.append((6, 0, 0))
            if nsh >= 2 and quality > 0:
                out.append((6, 0, 1))
    else:
        if (m >> 7) & 1:
            out.append((7, s2, 0))
        if (m >> 4) & 1:
            out += [(4, c, 0) for c in range(8)]
        if (m >> 5) & 1:
            out += [(5, c, 0) for c in range(4)]
        if (m >> 6) & 1:
            out.append((6, 0, 0))
    return out


def _logits(fh, rk, blocks):
    import torch
    n = blocks.shape[0]
    d_blk = torch.from_numpy(np.ascontiguousarray(blocks).view(np.int32)).cuda()
    d_log = torch.zeros((n, 512), dtype=torch.float32, device="cuda")
    rk.forward(d_blk.data_ptr(), n, d_log.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return d_log.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("quality,threshold,topk,modes", [(0, 0.0, 4, 0xFF),
                                                           (0, 2.0, 4, 0xFF),
                                                           (8, None, 3, 0xFF),
                                                           (5, 0.0, 2, 0xCF)])
def test_learned_search_matches_oracle(fh, oracle, quality, threshold, topk, modes):
    img = synth.generate(64, 64, 3)
    blocks = synth.blocks(img)
    rk = fh.Ranker(seed=11)
    z = _logits(fh, rk, blocks)
    valid = oracle.mlp_valid_mask() & np.repeat([(modes >> m) & 1 == 1 for m in range(8)], 64)
    zz = np.where(valid, z, -np.inf)
    pmax = 1.0 / np.exp(zz - zz.max(1, keepdims=True)).sum(1)
    if threshold is None:
        threshold = float(np.median(pmax))
    print(f"learned q={quality} thr={threshold:.4f} k={topk}", flush=True)
    got = fh.compress(img, quality=quality, seed=21, block_modes=modes,
                      wm_counter=ctypes.c_uint32(0xFFFFFFFF), ranker=rk, topk=topk,
                      threshold=threshold).reshape(-1, 16)
    plain, _ = oracle.bc7_compress(img, quality=0)
    plain = plain.reshape(-1, 16)
    checked = 0
    for b in range(blocks.shape[0]):
        px = blocks[b]
        if (px == px[0]).all() or not ((px >> 24) != 0).any():
            assert np.array_equal(got[b], plain[b]), b  # solid / transparent
            continue
        if abs(pmax[b] - threshold) < 1e-4:
            continue  # float32 on the GPU could decide either way
        if pmax[b] >= threshold:
            order = sorted(np.nonzero(valid)[0], key=lambda j: (-z[b, j], j))[:topk]
            cands = [(int(j) >> 6, int(j) & 63, 0) for j in order]
            want = oracle.bc7_encode_learned(px, cands, quality=quality, seed=21, block_id=b)
        else:
            cands = "full search"
            want, _ = oracle.bc7_compress_block(px, quality=quality, seed=21, block_id=b,
                                                block_modes=modes, q0_guard=True)
        assert np.array_equal(got[b], want), (b, cands)
        checked += 1
    assert checked > 150


@pytest.mark.parametrize("quality", [0, 9])
def test_learned_without_confident_blocks_is_the_full_search(fh, quality):
    """With a threshold no softmax probability reaches, every block falls back
    to the full search: the output is the full search's, byte for byte."""
    img = synth.generate(128, 128, 5)
    rk = fh.Ranker(seed=4)
    full = fh.compress(img, quality=quality, seed=8, wm_counter=ctypes.c_uint32(0xFFFFFFFF))
    got = fh.compress(img, quality=quality, seed=8, wm_counter=ctypes.c_uint32(0xFFFFFFFF),
                      ranker=rk, threshold=2.0)
    assert np.array_equal(got, full)


def test_learned_search_never_worse_than_its_quality0(fh, oracle):
    img = synth.generate(128, 128, 4)
    rk = fh.Ranker(seed=3)
    q0 = fh.compress(img, quality=0, ranker=rk, threshold=0.0, wm_counter=ctypes.c_uint32(0))
    q9 = fh.compress(img, quality=9, seed=2, ranker=rk, threshold=0.0,
                     wm_counter=ctypes.c_uint32(0))
    blocks = synth.blocks(img)
    d0 = synth.blocks(fh.decompress(q0, 128, 128))
    d9 = synth.blocks(fh.decompress(q9, 128, 128))
    for b in range(blocks.shape[0]):
        assert oracle.block_premul_err(blocks[b], d9[b]) <= \
            oracle.block_premul_err(blocks[b], d0[b])
