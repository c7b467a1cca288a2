"""BC7 parity: HIP encoder vs the CPU oracle on generated images.

The bar is bit-exact output: at quality 0 the search is deterministic; at
quality > 0 both sides use the same counter-based annealing stream.
"""
import ctypes

import numpy as np
import pytest

from fastc_amd import synth

pytestmark = pytest.mark.gpu


def _first_diff(a, b):
    a = a.reshape(-1, 16)
    b = b.reshape(-1, 16)
    bad = np.nonzero((a != b).any(axis=1))[0]
    return bad


@pytest.mark.parametrize("size,seed", [(64, 1), (128, 7)])
def test_bc7_q0_bit_exact(fh, oracle, size, seed):
    img = synth.generate(size, size, seed)
    print(f"bc7 q0 size={size} seed={seed}", flush=True)
    ref, ref_wm = oracle.bc7_compress(img, quality=0)
    wm = ctypes.c_uint32(0xFFFFFFFF)
    got = fh.compress(img, quality=0, wm_counter=wm)
    bad = _first_diff(got, ref)
    assert bad.size == 0, f"{bad.size} blocks differ, first {bad[:8]}"
    assert wm.value == ref_wm


def test_bc7_q50_counter_rng_bit_exact(fh, oracle):
    img = synth.generate(64, 64, 3)
    print("bc7 q50 64x64 seed=3", flush=True)
    ref, _ = oracle.bc7_compress(img, quality=50, seed=1234)
    got = fh.compress(img, quality=50, seed=1234, wm_counter=ctypes.c_uint32(0xFFFFFFFF))
    bad = _first_diff(got, ref)
    assert bad.size == 0, f"{bad.size} blocks differ, first {bad[:8]}"


def test_bc7_decode_matches_oracle(fh, oracle):
    img = synth.generate(64, 64, 5)
    data, _ = oracle.bc7_compress(img, quality=0)
    dec_ref = oracle.bc7_decompress(data, 64, 64)
    dec = fh.decompress(data, 64, 64)
    assert np.array_equal(dec, dec_ref)
