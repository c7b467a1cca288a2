"""The reference-built DXT/ETC1 library reproduces the committed golden
fixtures (so fixtures and live reference agree)."""
import numpy as np

from tests.ref_lib import ref


def test_ref_matches_golden():
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dxt_etc.npz"))
    r = ref()
    assert np.array_equal(r.dxt(g["image"], 0), g["dxt1"])
    assert np.array_equal(r.dxt(g["image"], 1), g["dxt5"])
    assert np.array_equal(r.etc1(g["image"]), g["etc1"])
