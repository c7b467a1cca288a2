"""Pins of the oracle's nonuniform error metric (m_ErrorMetric,
BPTCCompressor.h:73-77; weights kErrorMetrics, Compressor.cpp:205-208).

Each check follows from the definition of the weighted error, not from the
oracle's code:
* QuantizedError with equal endpoints is the weighted squared distance of
  every pixel to the endpoint colour (RGBAEndpoints.cpp:226-250);
* scaling every weight by 2 scales every error by exactly 4 and keeps every
  index (the weight multiplies the distance before squaring);
* the internal error of a block written in a mode without rotation equals
  the weighted squared distance of the decoded block to the source;
* the rotation modes weigh the rotated channel as alpha and alpha as the
  rotated channel (GetErrorMetric, CompressionMode.h:196-205), both in the
  colour fit and in the alpha term (Compressor.cpp:712);
* CompressSingleColor chooses its p-bit combination by the unrotated
  weights (Compressor.cpp:334-339), against exhaustive enumeration;
* BoxSelection's shape choice is the first strict minimum of the weighted
  estimates, against a rescan built on the pinned QuantizedError.
"""
import ctypes

import numpy as np
import pytest

from tests.test_oracle_bc7 import _MODES, _W1, _expand, _pbits_equal_per_subset
from tests.test_oracle_fit import _channels, _members, tables  # noqa: F401

W_NU = np.sqrt(np.array([0.3, 0.56, 0.11], np.float32)).astype(np.float64).tolist() + [1.0]


def _rand_block(rng, lo=0, hi=256, alpha=None):
    ch = rng.integers(lo, hi, size=(16, 4)).astype(np.uint32)
    if alpha is not None:
        ch[:, 3] = alpha
    return (ch[:, 0] | (ch[:, 1] << 8) | (ch[:, 2] << 16) | (ch[:, 3] << 24)).astype(np.uint32)


def test_metric_weights(oracle):
    w = oracle.metric_weights(1)
    assert np.allclose(np.square(w[:3]), [0.3, 0.56, 0.11], rtol=1e-6) and w[3] == 1.0
    assert oracle.metric_weights(0) == [1.0, 1.0, 1.0, 1.0]


def test_equal_endpoints_is_the_weighted_distance(oracle):
    rng = np.random.default_rng(11)
    weights = [W_NU, [1.0, 0, 0, 0], [0, 1.0, 0, 0], [0, 0, 1.0, 0], [0, 0, 0, 1.0],
               [0.25, 3.0, 0.5, 1.5]]
    for _ in range(40):
        px = _rand_block(rng)
        c = rng.integers(0, 256, 4).astype(np.float32)
        for w in weights:
            for nb in (4, 8, 16):
                e, idx = oracle.quantized_error(px, c, c, nb, weights=w)
                d = np.array([_channels(p) for p in px]) - c.astype(np.float64)
                want = float(((np.abs(d) * np.array(w, np.float32)) ** 2).sum())
                assert e == pytest.approx(want, rel=2e-6, abs=1e-9), (w, nb)
                assert not idx.any()


def test_weights_scale_the_error_squared(oracle):
    """Weights 2 (and 0.5) give exactly 4 (0.25) times the uniform error and
    the same indices, for random endpoints, index widths and masks; weighting
    the squared error instead would give a factor of 2."""
    rng = np.random.default_rng(12)
    for _ in range(60):
        px = _rand_block(rng)
        p1 = rng.integers(0, 256, 4).astype(np.float32)
        p2 = rng.integers(0, 256, 4).astype(np.float32)
        for nb in (4, 8, 16):
            for mask in (0xFFFFFFFF, 0xFCFCFCFC, 0xF8F8F8F8):
                e1, i1 = oracle.quantized_error(px, p1, p2, nb, mask)
                for f, sq in ((2.0, 4.0), (0.5, 0.25)):
                    ef, i_f = oracle.quantized_error(px, p1, p2, nb, mask, weights=[f] * 4)
                    assert ef == sq * e1 and np.array_equal(i_f, i1)


def _weighted_decoded_err(px, dec, w):
    a = np.array([_channels(p) for p in px])
    b = np.array([_channels(p) for p in dec])
    return float(((np.abs(a - b) * np.array(w)) ** 2).sum())


@pytest.mark.parametrize("quality", [0, 8])
def test_internal_error_is_the_weighted_decoded_error(oracle, quality):
    """Mode 1, and modes 0, 3, 6 and 7 where both endpoints of every subset
    carry the same p-bit (Pack's endpoint swap leaves the p-bits in place),
    encode exactly what their internal error measures, so under the
    nonuniform metric the chosen mode's error is the weighted squared
    distance of the decoded block (float accumulation: relative 1e-5)."""
    rng = np.random.default_rng(13)
    seen = set()
    for t in range(300):
        if t % 3 == 0:
            px = _rand_block(rng, 40, 200, alpha=255)
        elif t % 3 == 1:
            base = rng.integers(0, 256, 4)
            ch = np.clip(base + rng.integers(-20, 21, size=(16, 4)), 0, 255).astype(np.uint32)
            ch[:, 3] = 255
            px = (ch[:, 0] | (ch[:, 1] << 8) | (ch[:, 2] << 16) | (ch[:, 3] << 24))
        else:
            px = _rand_block(rng, 0, 256, alpha=rng.integers(0, 256, 16))
        px = np.ascontiguousarray(px, dtype=np.uint32)
        out, st = oracle.bc7_compress_block(px, quality=quality, seed=5, block_id=t, metric=1)
        if st.mode not in (0, 1, 3, 6, 7):
            continue
        if st.mode != 1 and not _pbits_equal_per_subset(out, st.mode):
            continue
        dec = oracle.bc7_decode_block_spec(out)
        want = _weighted_decoded_err(px, dec, W_NU)
        assert st.best_error == pytest.approx(want, rel=1e-5, abs=1e-6), (t, st.mode)
        seen.add(st.mode)
    assert {1, 3, 6, 7} <= seen, seen


def test_rotation_swaps_the_weights(oracle):
    """Rotation r (1..3) swaps channel r-1 with alpha before the fit.  With
    alpha in {0, v}, channel r-1 equal to alpha and the other channels 0 (v
    = 254 for mode 5, 251 for mode 4: values the colour grid holds under the
    zero p-bit QuantizedError applies), the rotated colour fit is exact in
    every colour channel and its error is alpha's distance to the fit's
    alpha, weighted by the rotated alpha weight w[r-1].  With channel r-1
    constant at 253, mode 4's alpha term is the equal-alpha table's error,
    weighted by the same w[r-1]."""
    for mode, v in ((5, 254), (4, 251)):
        for rot in (1, 2, 3):
            a = np.array([0, v] * 8, np.uint32)
            ch = np.zeros((16, 4), np.uint32)
            ch[:, 3] = a
            ch[:, rot - 1] = a
            px = (ch[:, 0] | (ch[:, 1] << 8) | (ch[:, 2] << 16) | (ch[:, 3] << 24))
            _, _, _, _, _, tu = oracle.fit_alpha(px, mode, rot, 0, metric=0)
            _, _, _, _, _, tn = oracle.fit_alpha(px, mode, rot, 0, metric=1)
            if mode == 5:  # the fit's alpha is 255: eight pixels 255 away, eight 1
                assert tu.rgb_error == 8 * 255.0 ** 2 + 8
            assert tn.rgb_error == pytest.approx(W_NU[rot - 1] ** 2 * tu.rgb_error, rel=1e-6)
            if mode == 4:
                rng = np.random.default_rng(rot)
                ch = rng.integers(0, 256, size=(16, 4)).astype(np.uint32)
                ch[:, rot - 1] = 253
                px = (ch[:, 0] | (ch[:, 1] << 8) | (ch[:, 2] << 16) | (ch[:, 3] << 24))
                _, _, _, _, _, tu = oracle.fit_alpha(px, 4, rot, 0, metric=0)
                _, _, _, _, _, tn = oracle.fit_alpha(px, 4, rot, 0, metric=1)
                assert tu.alpha_error > 0
                assert tn.alpha_error == pytest.approx(W_NU[rot - 1] ** 2 * tu.alpha_error,
                                                       rel=1e-6)


def test_single_color_in_rotation_uses_unrotated_weights(oracle):
    """When only channel r-1 varies, the rotated points are one colour and
    the colour fit is CompressSingleColor's, which weighs by the unrotated
    metric (Compressor.cpp:334): alpha (1, off mode 4's single-colour grid)
    now sits in channel r-1, so the error scales by w[r-1]^2, not by the
    rotated weight 1."""
    rng = np.random.default_rng(9)
    for rot in (1, 2, 3):
        ch = np.zeros((16, 4), np.uint32)
        ch[:, 3] = 1
        ch[:, rot - 1] = rng.integers(0, 256, 16)
        px = (ch[:, 0] | (ch[:, 1] << 8) | (ch[:, 2] << 16) | (ch[:, 3] << 24))
        _, _, _, _, _, tu = oracle.fit_alpha(px, 4, rot, 0, metric=0)
        _, _, _, _, _, tn = oracle.fit_alpha(px, 4, rot, 0, metric=1)
        assert tu.path in (oracle.FIT_SAME_POINT, oracle.FIT_ONE_BUCKET) and tu.rgb_error == 16.0
        assert tn.rgb_error == pytest.approx(W_NU[rot - 1] ** 2 * 16.0, rel=1e-6)


def test_alpha_term_weight_without_rotation(oracle):
    """Rotation 0 keeps alpha's weight 1: the alpha term is unchanged."""
    rng = np.random.default_rng(14)
    for _ in range(20):
        px = _rand_block(rng)
        for mode in (4, 5):
            _, _, _, _, _, tu = oracle.fit_alpha(px, mode, 0, 0, metric=0)
            _, _, _, _, _, tn = oracle.fit_alpha(px, mode, 0, 0, metric=1)
            assert tn.alpha_error == tu.alpha_error


def test_single_color_nonuniform_brute_force(oracle):
    """The p-bit combination minimises the sum of (distance * weight)^2 with
    the unrotated weights, accumulated in float32 in channel order (first
    combination on ties)."""
    L = oracle.lib()
    rng = np.random.default_rng(15)
    cols = [[0, 0, 0, 0], [255, 255, 255, 255], [1, 128, 254, 7]] + \
        [list(rng.integers(0, 256, 4)) for _ in range(40)]
    w32 = np.array(W_NU, np.float32)
    for mode, (cb, ab, pbt, ib, aib) in _MODES.items():
        for im in ([0, 1] if mode == 4 else [0]):
            w = _W1[ib if im == 0 else aib]
            combos = [(0, 0)] if pbt == 0 else ([(0, 0), (1, 1)] if pbt == 1 else
                                                [(0, 0), (0, 1), (1, 0), (1, 1)])
            for col in cols:
                p = np.array(col, np.float32)
                p1 = np.zeros(4, np.float32)
                p2 = np.zeros(4, np.float32)
                pbc = ctypes.c_int(0)
                err = L.ora_single_color_m(mode, im, 1, p.ctypes.data, p1.ctypes.data,
                                           p2.ctypes.data, ctypes.byref(pbc))
                best, best_c = None, None
                for ci, (pl, ph) in enumerate(combos):
                    e = np.float32(0.0)
                    for chn in range(4):
                        bits = cb if chn < 3 else ab
                        val = int(col[chn])
                        if bits == 0:
                            d = 255 - val
                        else:
                            pbl, pbh = (pl, ph) if pbt else (-1, -1)
                            lo = np.array([_expand(k, bits, pbl) for k in range(1 << bits)])
                            hi = np.array([_expand(k, bits, pbh) for k in range(1 << bits)])
                            ip = ((64 - w) * lo[:, None] + w * hi[None, :] + 32) >> 6
                            d = int(np.abs(ip - val).min())
                        x = np.float32(d) * w32[chn]
                        e = np.float32(e + np.float32(x * x))
                    if best is None or e < best:
                        best, best_c = e, ci
                assert err == float(best) and pbc.value == best_c, (mode, im, col)


def _estimate_w(oracle, px, members, nb, w):
    pts = np.array([_channels(px[i]) for i in members])
    lo, hi = pts.min(axis=0), pts.max(axis=0)
    if ((hi - lo) ** 2).sum() == 0:
        return 0.0
    e, _ = oracle.quantized_error(px, lo, hi, nb, weights=w, sel=members)
    return 0.0001 + e


def test_box_selection_nonuniform_rescan(oracle, tables):  # noqa: F811
    """BoxSelection under the nonuniform metric picks the first strict
    minimum of the weighted estimates (Compressor.cpp:1670-1750)."""
    L = oracle.lib()
    rng = np.random.default_rng(16)
    differs = 0
    for t in range(40):
        alpha = 255 if t % 2 == 0 else rng.integers(0, 256, 16)
        px = np.ascontiguousarray(_rand_block(rng, 0, 256, alpha=alpha), dtype=np.uint32)
        if t % 4 == 0:  # few distinct colours: ties and early exits
            px = np.ascontiguousarray(px[rng.integers(0, 3, 16)], dtype=np.uint32)
        opaque = all((int(p) >> 24) >= 250 for p in px)
        best2 = best3 = None
        sh2 = sh3 = 0
        want = None
        for i in range(64):
            err = sum(_estimate_w(oracle, px, _members(tables, 2, i, s), 8, W_NU)
                      for s in range(2))
            if best2 is None or err < best2:
                best2, sh2 = err, i
            if err < 1e-9:
                want = (1, sh2, 0b10001010)
                break
        if want is None and not opaque:
            want = (1, sh2, 0b11110000)
        if want is None:
            for i in range(64):
                err = sum(_estimate_w(oracle, px, _members(tables, 3, i, s), 4, W_NU)
                          for s in range(3))
                if best3 is None or err < best3:
                    best3, sh3 = err, i
                if err < 1e-9:
                    want = (2, sh2, 0b00000101)
                    break
            if want is None:
                want = (2, sh2, 0b11001111)
        ns, s2, s3, modes = (ctypes.c_uint32() for _ in range(4))
        L.ora_bc7_box_selection_m(px.ctypes.data, 1, ctypes.byref(ns), ctypes.byref(s2),
                                  ctypes.byref(s3), ctypes.byref(modes))
        assert (ns.value, s2.value, modes.value) == want, t
        if want[0] == 2 and want[2] == 0b11001111:
            assert s3.value == sh3, t
        u2 = ctypes.c_uint32()
        L.ora_bc7_box_selection_m(px.ctypes.data, 0, ctypes.byref(ns), ctypes.byref(u2),
                                  ctypes.byref(s3), ctypes.byref(modes))
        differs += u2.value != s2.value
    assert differs > 0  # the metric changes some choices
