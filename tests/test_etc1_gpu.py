"""ETC1 parity: the HIP encoder against the reference's own rg_etc1 build
(ETCEncoder/src/rg_etc1.cpp at cLowQuality, as ETCEncoder/src/Compressor.cpp
configures it): committed golden output for a 64x64 image and SHA-256 of the
reference's output at 256x256 and 2048x2048."""
import hashlib
import json
import os

import numpy as np
import pytest

from fastc_amd import synth

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_etc1_golden_64(fh):
    g = np.load(os.path.join(GOLD, "dxt_etc.npz"))
    got = fh.compress(g["image"], fmt=2)
    bad = np.nonzero((got.reshape(-1, 8) != g["etc1"].reshape(-1, 8)).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} blocks differ, first {bad[:8]}"


@pytest.mark.parametrize("size", [256, 2048])
def test_etc1_full_size_hash(fh, size):
    print(f"etc1 {size}", flush=True)
    hashes = json.load(open(os.path.join(GOLD, "dxt_etc_hashes.json")))
    img = synth.generate(size, size, 2, edges=True)
    got = fh.compress(img, fmt=2)
    assert hashlib.sha256(got.tobytes()).hexdigest() == hashes[str(size)]["etc1"]


def test_etc1_solid_colours(fh):
    """Solid blocks of many colours, including 0 and 255 channels, take the
    table-driven solid path (rg_etc1.cpp:1951-2033)."""
    ref = pytest.importorskip("tests.ref_lib").ref()
    vals = [0, 1, 2, 7, 128, 200, 254, 255]
    cols = [(r | (g << 8) | (b << 16) | 0xFF000000) for r in vals for g in vals for b in vals]
    img = np.zeros((4, 4 * len(cols)), np.uint32)
    for i, c in enumerate(cols):
        img[:, 4 * i:4 * i + 4] = c
    got = fh.compress(img, fmt=2)
    want = ref.etc1(img)
    assert np.array_equal(got, want)
