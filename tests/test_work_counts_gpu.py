"""Work counting of the BC7 annealing (fh_set_work_counting,
fh_bc7_work_counts): the GPU's count of the reference's QuantizedError point
evaluations equals the oracle's for the same encode (the GPU is bit-exact
with the oracle, so both take the same annealing steps), the GPU's own
pixel scores are 16 per candidate step, and counting leaves the bytes
unchanged."""
import ctypes

import numpy as np
import pytest

from fastc_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w,h,seed,q", [(64, 64, 3, 50), (128, 64, 5, 12)])
def test_anneal_point_count_matches_oracle(fh, oracle, w, h, seed, q):
    img = synth.generate(w, h, seed)
    plain = fh.compress(img, quality=q, seed=1234, wm_counter=ctypes.c_uint32(0xFFFFFFFF))
    fh.set_work_counting(True)
    try:
        got = fh.compress(img, quality=q, seed=1234, wm_counter=ctypes.c_uint32(0xFFFFFFFF))
        n_ref, n_gpu = fh.work_counts()
    finally:
        fh.set_work_counting(False)
    assert np.array_equal(got, plain)
    L = oracle.lib()
    L.ora_anneal_point_count(1)
    oracle.bc7_compress(img, quality=q, seed=1234, q0_guard=True)
    want = L.ora_anneal_point_count(1)
    print(f"{w}x{h} q={q}: reference point evaluations {n_ref} (oracle {want}), "
          f"GPU pixel scores {n_gpu}", flush=True)
    assert want > 0 and n_ref == want
    assert n_gpu % 16 == 0 and n_gpu >= n_ref


def test_learned_search_counts(fh):
    """The learned search anneals fewer candidates than the full search."""
    import os
    weights = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                           "fastc_amd", "data", "bc7_ranker.npy")
    img = synth.generate(128, 128, 7)
    rk = fh.Ranker()
    rk.set_params(np.load(weights))
    fh.set_work_counting(True)
    try:
        fh.compress(img, quality=50, seed=3)
        full = fh.work_counts()
        fh.compress(img, quality=50, seed=3, ranker=rk)
        learned = fh.work_counts()
    finally:
        fh.set_work_counting(False)
    print(f"full {full}, learned {learned}", flush=True)
    assert 0 < learned[0] < full[0]
    assert learned[1] % 16 == 0 and learned[1] >= learned[0]
