"""Training labels of the candidate ranker: the exhaustive quality-0 GPU
search (fh_bc7_exhaustive_labels) against the oracle's, bit for bit."""
import numpy as np
import pytest

from fastc_amd import synth

pytestmark = pytest.mark.gpu


def test_exhaustive_labels_match_oracle(fh, oracle):
    import torch
    img = synth.generate(64, 64, 5)
    blocks = synth.blocks(img)
    # drop solid and fully transparent blocks: they never reach the search
    keep = ~(blocks == blocks[:, :1]).all(axis=1) & ((blocks >> 24) != 0).any(axis=1)
    blocks = blocks[keep]
    n = blocks.shape[0]
    print(f"labels n={n}", flush=True)
    d_blk = torch.from_numpy(blocks.view(np.int32).copy()).cuda()
    d_lab = torch.zeros(n, dtype=torch.int32, device="cuda")
    d_err = torch.zeros(n, dtype=torch.int64, device="cuda")
    fh.exhaustive_labels_device(d_blk.data_ptr(), n, d_lab.data_ptr(), d_err.data_ptr(),
                                torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    lab = d_lab.cpu().numpy()
    err = d_err.cpu().numpy().view(np.uint64)
    for i in range(n):
        want_lab, want_err = oracle.bc7_exhaustive_label(blocks[i])
        assert (lab[i], int(err[i])) == (want_lab, want_err), i
    # the exhaustive search is never worse than the reference search
    worse = 0
    for i in range(n):
        o, _ = oracle.bc7_compress_block(blocks[i])
        d = oracle.bc7_decompress(o, 4, 4).reshape(16)
        worse += int(err[i]) > oracle.block_premul_err(blocks[i], d)
    assert worse == 0
