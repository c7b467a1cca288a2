"""Stage-by-stage pins of the BC7 oracle's cluster fits, BoxSelection and Pack
(no GPU).

The reference BC7 encoder cannot be built here (its headers need CMake-made
config files), so its float trajectories cannot be replayed.  Each stage of
CompressCluster (Compressor.cpp:632-1094) is pinned instead by what the
mathematics fixes at that stage, checked on the oracle's own intermediate
values (ora_fit_trace):

* GetPrincipalAxis (RGBAEndpoints.cpp:327-428): the float64 dominant
  eigenvector of the covariance, signed like the power method's start
  vector, and hand-derived known answers of the collinear shortcut, including
  the read of (-1,-1,-1,-1) padding past the unique points (:364-371);
* the Lloyd k-means (:973-1026) at termination: a fixed point, every point in
  its first strictly nearest centroid, every centroid the exact mean of its
  bucket and every empty one at the origin;
* the least-squares endpoints (:1053-1077): the float64 solution of the 2x2
  normal equations of the final buckets;
* ClampEndpointsToGrid (:212-250): a float64 brute force over the p-bit
  combinations, first strict minimum;
* whole fits: two grid colours per subset are reproduced exactly (error 0,
  decoded bit for bit);
* the rotation fit (:632-919): known answers of the stale-statistics quirk
  and of the equal-alpha mode-4 table branch, exact two-level alpha;
* BoxSelection (:1670-1750): a rescan of every shape from the definitions on
  the pinned QuantizedError, checking the chosen shape indices;
* Pack (:1096-1298): an independent reader of the BC7 specification decodes
  packed parameters of every mode back to their pixels, and agrees with the
  oracle's specification decoder on random blocks.
"""
import ctypes

import numpy as np
import pytest

from fastc_amd import synth

# (colour bits, alpha bits, p-bit type 0 none / 1 shared / 2 per endpoint,
#  colour index bits, alpha index bits, subsets, partition bits) (BC7 spec)
SPEC = {0: (4, 0, 2, 3, 0, 3, 4), 1: (6, 0, 1, 3, 0, 2, 6), 2: (5, 0, 0, 2, 0, 3, 6),
        3: (7, 0, 2, 2, 0, 2, 6), 4: (5, 6, 0, 2, 3, 1, 0), 5: (7, 8, 0, 2, 2, 1, 0),
        6: (7, 7, 2, 4, 0, 1, 0), 7: (5, 5, 2, 2, 0, 2, 6)}
# interpolation weights of 2-, 3- and 4-bit indices (BC7 spec)
WEIGHTS = {2: [0, 21, 43, 64], 3: [0, 9, 18, 27, 37, 46, 55, 64],
           4: [0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64]}
START = np.full(4, 0.5)


def _combos(pbt):
    return [(-1, -1)] if pbt == 0 else ([(0, 0), (1, 1)] if pbt == 1 else
                                        [(0, 0), (0, 1), (1, 0), (1, 1)])


def _mask(mode):
    """Per-channel quantisation mask: the top `bits` bits of each channel;
    a mode without alpha has an all-zero alpha byte."""
    cb, ab = SPEC[mode][0], SPEC[mode][1]
    m = 0
    for ch in range(3):
        m |= ((0xFF << (8 - cb)) & 0xFF) << (8 * ch)
    if ab:
        m |= ((0xFF << (8 - ab)) & 0xFF) << 24
    return m


def _expand(k, bits, pbit):
    """Endpoint value k of `bits` bits (p-bit appended when >= 0), widened to
    8 bits by replicating its top bits (BC7 spec)."""
    n = bits + (pbit >= 0)
    v = (k << 1 | pbit) if pbit >= 0 else k
    v <<= 8 - n
    return (v | (v >> n)) & 0xFF


@pytest.fixture(scope="module")
def tables(oracle):
    L = oracle.lib()
    s2 = np.zeros(64, np.uint16)
    s3 = np.zeros(128, np.uint16)
    a2 = np.zeros(64, np.uint8)
    a3 = np.zeros(128, np.uint8)
    iv = np.zeros(128, np.uint32)
    L.ora_table_partitions(s2.ctypes.data, s3.ctypes.data, a2.ctypes.data, a3.ctypes.data,
                           iv.ctypes.data)
    return s2.astype(int), s3.astype(int), a2.astype(int), a3.astype(int)


def _subset_of(tables, nsub, shape, i):
    s2, s3 = tables[0], tables[1]
    if nsub == 2:
        return (s2[shape] >> i) & 1
    if nsub == 3:
        if (s3[2 * shape] >> i) & 1:
            return 1 + ((s3[2 * shape + 1] >> i) & 1)
        return 0
    return 0


def _anchor(tables, nsub, shape, s):
    a2, a3 = tables[2], tables[3]
    if s == 0:
        return 0
    if nsub == 2:
        return a2[shape]
    return a3[shape] if s == 1 else a3[64 + shape]


def _members(tables, nsub, shape, subset):
    return [i for i in range(16) if _subset_of(tables, nsub, shape, i) == subset]


def _channels(p):
    return np.array([(int(p) >> (8 * k)) & 0xFF for k in range(4)], dtype=np.float64)


def _to_pixel(oracle, v, mask, pbit):
    v = np.ascontiguousarray(v, dtype=np.float32)
    return oracle.lib().ora_to_pixel(v.ctypes.data, mask, pbit)


def _grid_brute_force(oracle, mode, p1, p2):
    """ClampEndpointsToGrid from its definition: clamp to [0, 255], quantise
    both endpoints under each p-bit combination (the pinned ToPixel), keep
    the first combination with the strictly smallest float64 squared
    distance.  Returns (combination, q1, q2, sorted distances)."""
    mask = _mask(mode)
    c1 = np.clip(np.asarray(p1, np.float32), 0, 255)
    c2 = np.clip(np.asarray(p2, np.float32), 0, 255)
    best, dists = None, []
    for ci, (pa, pb) in enumerate(_combos(SPEC[mode][2])):
        q1 = _channels(_to_pixel(oracle, c1, mask, pa))
        q2 = _channels(_to_pixel(oracle, c2, mask, pb))
        d = float(((q1 - c1.astype(np.float64)) ** 2).sum() +
                  ((q2 - c2.astype(np.float64)) ** 2).sum())
        dists.append(d)
        if best is None or d < best[3]:
            best = (ci, q1, q2, d)
    return best[0], best[1], best[2], sorted(dists)


# --------------------------------------------------------------------------
# ClampEndpointsToGrid
# --------------------------------------------------------------------------

@pytest.mark.parametrize("mode", range(8))
def test_clamp_to_grid_brute_force(oracle, mode):
    """Integer endpoints give exact float distances, so ties between p-bit
    combinations occur and must go to the first combination; fractional and
    out-of-range endpoints exercise the clamp and the rounding."""
    rng = np.random.default_rng(100 + mode)
    ties = 0
    for trial in range(600):
        if trial % 6 == 0:
            # two even and two odd channels per endpoint: with a p-bit as the
            # lowest bit, both p-bits of an endpoint cost the same
            par = np.array([0, 0, 1, 1])
            p1 = (rng.integers(0, 128, 4) * 2 + rng.permutation(par)).astype(np.float32)
            p2 = (rng.integers(0, 128, 4) * 2 + rng.permutation(par)).astype(np.float32)
        elif trial % 3 == 0:
            p1 = rng.integers(0, 256, 4).astype(np.float32)
            p2 = rng.integers(0, 256, 4).astype(np.float32)
        else:
            p1 = rng.uniform(-30, 290, 4).astype(np.float32)
            p2 = rng.uniform(-30, 290, 4).astype(np.float32)
        want_c, q1, q2, d = _grid_brute_force(oracle, mode, p1, p2)
        g1, g2, got_c = oracle.clamp_to_grid(mode, p1, p2)
        exact = trial % 3 == 0
        if len(d) > 1 and d[1] == d[0]:
            ties += 1
        if not exact and len(d) > 1 and d[1] - d[0] <= 1e-3 * max(d[0], 1.0):
            continue  # a float32 near-tie: either choice is the reference's
        assert got_c == want_c, (mode, p1, p2, d)
        assert np.array_equal(g1, q1) and np.array_equal(g2, q2), (mode, p1, p2)
    if mode in (0, 1, 6, 7):
        # the tie-break was exercised (mode 3's three channels cannot tie at
        # the minimum: one p-bit always fits more of them)
        assert ties > 0


# --------------------------------------------------------------------------
# GetPrincipalAxis
# --------------------------------------------------------------------------

def _pix(r, g, b, a):
    return (int(r) | int(g) << 8 | int(b) << 16 | int(a) << 24)


def test_principal_axis_known_answers(oracle):
    """Hand-derived answers of GetPrincipalAxis (RGBAEndpoints.cpp:327-428)."""
    # 16 distinct collinear points, the first two out of order: the collinear
    # shortcut returns normalize(p1 - p0) = -R (the power method would give +R)
    xs = [160, 10] + list(range(20, 160, 10))
    px = np.array([_pix(x, 20, 30, 255) for x in xs], np.uint32)
    assert np.array_equal(oracle.principal_axis(px), [-1, 0, 0, 0])
    # one duplicate: the loop over all 16 points reads an unused unique-point
    # slot, (-1,-1,-1,-1), off the line -> not collinear -> covariance path;
    # only the RR term is non-zero, so the power method from (1/2,..,1/2)
    # returns +R after one product
    xs2 = xs[:15] + [xs[14]]
    px = np.array([_pix(x, 20, 30, 255) for x in xs2], np.uint32)
    assert np.array_equal(oracle.principal_axis(px), [1, 0, 0, 0])
    # points on the grey diagonal through the origin, with duplicates: the
    # padding (-1,-1,-1,-1) lies on the line, so the shortcut is taken and the
    # axis is normalize(p1 - p0) = -(1/2,1/2,1/2,1/2)
    ks = [40, 0, 10, 20, 30, 10, 20, 30, 40, 0, 10, 20, 30, 10, 20, 30]
    px = np.array([_pix(k, k, k, k) for k in ks], np.uint32)
    assert np.array_equal(oracle.principal_axis(px), [-0.5, -0.5, -0.5, -0.5])
    # the same line shifted off the origin: the padding is off the line ->
    # covariance path -> +(1/2,1/2,1/2,1/2)
    px = np.array([_pix(k + 5, k + 5, k + 5, k + 9) for k in ks], np.uint32)
    ax = oracle.principal_axis(px)
    assert np.allclose(ax, [0.5, 0.5, 0.5, 0.5], atol=1e-6), ax
    # a single unique point: the zero axis
    px = np.full(16, _pix(7, 8, 9, 10), np.uint32)
    assert not oracle.principal_axis(px).any()
    # two unique points, all distinct from each other only in G: the padding
    # makes the covariance path run; the axis is +G whatever the order
    px = np.array([_pix(3, 90 if i % 2 else 10, 3, 255) for i in range(16)], np.uint32)
    assert np.allclose(oracle.principal_axis(px), [0, 1, 0, 0], atol=1e-6)
    px = np.array([_pix(3, 10 if i % 2 else 90, 3, 255) for i in range(16)], np.uint32)
    assert np.allclose(oracle.principal_axis(px), [0, 1, 0, 0], atol=1e-6)


def _cov64(pts):
    pts = np.asarray(pts, np.float64)
    d = pts - pts.mean(axis=0)
    return d.T @ d / 3.0


def _power4(m):
    b = START.copy()
    for _ in range(4):
        nb = m @ b
        n = np.linalg.norm(nb)
        if n < 1e-10:
            break
        nb /= n
        done = abs(1 - b @ nb) < 1e-8
        b = nb
        if done:
            break
    return b


def _random_blocks(rng, n):
    img = synth.generate(128, 128, 17)
    gen = [img[y:y + 4, x:x + 4].reshape(16).copy() for y in range(0, 128, 4)
           for x in range(0, 128, 4)]
    out = [b for b in gen if not (b == b[0]).all()][:n // 2]
    while len(out) < n:
        k = len(out) % 3
        if k == 0:
            b = rng.integers(0, 2**32, 16, dtype=np.uint64).astype(np.uint32)
        elif k == 1:
            b = (rng.integers(0, 2**24, 16) | 0xFF000000).astype(np.uint32)
        else:  # a noisy ramp: a dominant axis
            t = rng.integers(0, 200, 16)
            d = rng.integers(-1, 2, (16, 4))
            dirv = rng.integers(0, 3, 4)
            base = rng.integers(0, 40, 4)
            ch = np.clip(base + np.outer(t, dirv) // 3 + d, 0, 255)
            b = (ch[:, 0] | ch[:, 1] << 8 | ch[:, 2] << 16 | ch[:, 3] << 24).astype(np.uint32)
        out.append(b)
    return out


def test_principal_axis_is_the_dominant_eigenvector(oracle, tables):
    """Where the covariance's two largest eigenvalues separate (ratio < 0.05),
    four power steps leave the axis within 1e-3 of the float64 eigenvector,
    signed like the start vector (1/2,1/2,1/2,1/2); everywhere else the axis
    follows four float64 power steps of the same matrix."""
    rng = np.random.default_rng(7)
    blocks = _random_blocks(rng, 240)
    gap = general = 0
    for b in blocks:
        for nsub, shape in ((1, 0), (2, int(rng.integers(64))), (3, int(rng.integers(64)))):
            for subset in range(nsub):
                idx = _members(tables, nsub, shape, subset)
                pts = np.array([_channels(b[i]) for i in idx])
                uniq = np.unique(pts, axis=0)
                if len(uniq) < 3:
                    continue
                d = uniq - uniq[0]
                if np.linalg.matrix_rank(d, tol=1e-6) < 2:
                    continue  # collinear: the shortcut, pinned by known answers
                ax = oracle.principal_axis(b, nsub, shape, subset)
                m = _cov64(pts)
                w, v = np.linalg.eigh(m)
                v1 = v[:, -1]
                s = START @ v1
                if w[-1] > 0 and w[-2] / w[-1] < 0.05 and abs(s) > 0.1:
                    assert np.allclose(ax, np.sign(s) * v1, atol=1e-3), (ax, v1, w)
                    gap += 1
                assert np.allclose(ax, _power4(m), atol=2e-3), (ax, _power4(m))
                general += 1
    assert gap >= 50 and general >= 300, (gap, general)


# --------------------------------------------------------------------------
# k-means, least squares, grid: the RGB fit
# --------------------------------------------------------------------------

def _fits(oracle, tables, blocks, rng):
    for b in blocks:
        for mode in (0, 1, 2, 3, 6, 7):
            nsub = SPEC[mode][5]
            shape = int(rng.integers(16 if mode == 0 else 64)) if nsub > 1 else 0
            for subset in range(nsub):
                e, p1, p2, idx, pbc, tr = oracle.fit_rgb(b, mode, shape, subset)
                members = _members(tables, nsub, shape, subset)
                pts = np.array([_channels(b[i]) for i in members])
                yield mode, shape, subset, b, members, pts, e, p1, p2, idx, pbc, tr


def test_rgb_fit_stages(oracle, tables):
    """Every stage of CompressCluster's RGB fit (Compressor.cpp:921-1094) on
    several thousand fits of generator, random and ramp blocks."""
    rng = np.random.default_rng(11)
    blocks = _random_blocks(rng, 160)
    for k in range(20):  # flat halves: single-colour subsets
        c = rng.integers(0, 2**32, 2, dtype=np.uint64).astype(np.uint32)
        blocks.append(np.array([c[int(i >= 8)] for i in range(16)], np.uint32))
    for k in range(20):  # transparent black among a few colours: exact ties
        c = rng.integers(0, 2**32, 3, dtype=np.uint64).astype(np.uint32)
        c[0] = 0
        blocks.append(c[rng.integers(0, 3, 16)].astype(np.uint32))
    paths = {1: 0, 2: 0, 3: 0}
    ties_skipped = exact_ties = 0
    for mode, shape, subset, b, members, pts, e, p1, p2, idx, pbc, tr in \
            _fits(oracle, tables, blocks, rng):
        n = len(members)
        paths[tr.path] += 1
        mask = _mask(mode)
        if tr.path == oracle.FIT_SAME_POINT:
            assert (pts == pts[0]).all()
            continue
        axis = np.array(tr.axis, np.float64)
        # start endpoints: avg + {min, max}((p - avg) . axis) axis, clamped
        avg = pts.mean(axis=0)
        dp = (pts - avg) @ axis
        s1 = np.clip(avg + dp.min() * axis, 0, 255)
        s2 = np.clip(avg + dp.max() * axis, 0, 255)
        assert np.allclose(tr.start1, s1, atol=2e-3) and np.allclose(tr.start2, s2, atol=2e-3)
        nb = 1 << SPEC[mode][3]
        assert tr.nbuckets == nb and tr.iterations >= 1
        cent = np.array(tr.centroid, np.float64)[:nb]
        count = np.array(tr.count)[:nb]
        bucket = np.array(tr.bucket)[:n]
        assert np.array_equal(np.bincount(bucket, minlength=nb), count)
        for j in range(nb):
            if count[j]:
                # sums of integer channels are exact in float32: the centroid
                # is the correctly rounded mean
                want = (pts[bucket == j].sum(axis=0).astype(np.float32) /
                        np.float32(count[j]))
                assert np.array_equal(cent[j].astype(np.float32), want), (j, cent[j], want)
            else:
                assert not cent[j].any()  # an empty bucket sits at the origin
        # fixed point: each point in its first strictly nearest centroid
        for i in range(n):
            d = ((pts[i] - cent) ** 2).sum(axis=1)
            order = np.argsort(d, kind="stable")
            if d[order[1]] == d[order[0]] and (cent == np.round(cent)).all():
                # an exact tie (integer centroids: float32 distances are
                # exact too) goes to the first bucket
                exact_ties += 1
                assert bucket[i] == order[0], (i, d, bucket[i])
                continue
            if d[order[1]] - d[order[0]] <= 1e-4 * max(d[order[0]], 1.0):
                ties_skipped += 1
                assert d[bucket[i]] <= d[order[0]] * (1 + 1e-4) + 1e-4
                continue
            assert bucket[i] == order[0], (i, d, bucket[i])
        if tr.path == oracle.FIT_ONE_BUCKET:
            assert (count > 0).sum() == 1
            j = int(np.flatnonzero(count)[0])
            assert np.array_equal(np.array(tr.single, np.float32), cent[j].astype(np.float32))
            q1 = np.zeros(4, np.float32)
            q2 = np.zeros(4, np.float32)
            c = ctypes.c_int(0)
            single = np.array(tr.single, np.float32)
            e1 = oracle.lib().ora_single_color(mode, 0, single.ctypes.data, q1.ctypes.data,
                                               q2.ctypes.data, ctypes.byref(c))
            assert e == n * e1 and (idx[:n] == 1).all()
            continue
        # least squares over the final buckets, in float64
        a = (nb - 1 - np.arange(nb)) / (nb - 1)
        bw = np.arange(nb) / (nb - 1)
        asq, bsq, ab = (count * a * a).sum(), (count * bw * bw).sum(), (count * a * bw).sum()
        ax = (count * a) @ cent
        bx = (count * bw) @ cent
        det = asq * bsq - ab * ab
        l1 = (ax * bsq - bx * ab) / det
        l2 = (bx * asq - ax * ab) / det
        scale = (np.abs(ax * bsq) + np.abs(bx * ab) + np.abs(bx * asq) + np.abs(ax * ab)) / det
        tol = 2e-5 * scale + 1e-3
        assert np.all(np.abs(np.array(tr.lsq1) - l1) <= tol), (tr.lsq1, l1, tol)
        assert np.all(np.abs(np.array(tr.lsq2) - l2) <= tol), (tr.lsq2, l2, tol)
        # the grid clamp of the least-squares endpoints
        want_c, q1, q2, d = _grid_brute_force(oracle, mode, tr.lsq1, tr.lsq2)
        if len(d) == 1 or d[1] - d[0] > 1e-3 * max(d[0], 1.0):
            assert tr.pbit_combo == want_c
            assert np.array_equal(tr.grid1, q1) and np.array_equal(tr.grid2, q2)
        # quality 0: the error of the grid endpoints with QuantizedError
        combo = _combos(SPEC[mode][2])[tr.pbit_combo]
        pb = None
        if SPEC[mode][2]:
            pb = np.array(combo, np.int32)
        else:
            pb = np.zeros(2, np.int32)  # GetPBitCombo's {0, 0} (CompressionMode.h:243-250)
        sel = np.array(members, np.uint8)
        g1 = np.array(tr.grid1, np.float32)
        g2 = np.array(tr.grid2, np.float32)
        qidx = np.zeros(16, np.uint8)
        qe = oracle.lib().ora_quantized_error(b.ctypes.data, sel.ctypes.data, n, g1.ctypes.data,
                                              g2.ctypes.data, nb, mask, pb.ctypes.data,
                                              qidx.ctypes.data)
        assert e == qe and np.array_equal(idx[:n], qidx[:n])
        assert np.array_equal(p1, g1) and np.array_equal(p2, g2)
    # (a lone filled bucket after k-means is pinned by the stale-statistics
    # known answer below)
    assert paths[1] > 20 and paths[3] > 1500, paths
    assert ties_skipped < 200 and exact_ties > 0, (ties_skipped, exact_ties)


def test_kmeans_tie_goes_to_the_first_bucket(oracle):
    """Transparent black and opaque white in mode 6 (16 buckets): after the
    first pass buckets 1..14 are empty and sit at the origin, exactly where
    the black points are, so every black point ties between bucket 0 and
    those 14.  The first strict minimum keeps them in bucket 0, the least
    squares then land on the two colours (black with p-bit 0, white with
    p-bit 1) and the fit is exact."""
    px = np.array([_pix(255, 255, 255, 255) if i % 2 else 0 for i in range(16)], np.uint32)
    e, p1, p2, idx, pbc, tr = oracle.fit_rgb(px, 6)
    assert tr.path == oracle.FIT_LEAST_SQUARES
    assert list(tr.count) == [8] + [0] * 14 + [8]
    assert list(tr.bucket) == [0, 15] * 8
    assert np.array_equal(tr.grid1, [0, 0, 0, 0]) and np.array_equal(tr.grid2, [255] * 4)
    assert e == 0.0 and pbc == 1  # the combination (0, 1)


# --------------------------------------------------------------------------
# Whole fits: exact two-colour subsets
# --------------------------------------------------------------------------

def _grid_colour(rng, mode, pbit):
    cb, ab = SPEC[mode][0], SPEC[mode][1]
    c = [_expand(int(rng.integers(1 << cb)), cb, pbit) for _ in range(3)]
    a = _expand(int(rng.integers(1 << ab)), ab, pbit) if ab else 255
    return _pix(*c, a)


@pytest.mark.parametrize("mode", [0, 1, 3, 6, 7])
def test_two_grid_colours_per_subset_fit_exactly(oracle, tables, mode):
    """Subsets holding at most two colours of the mode's grid (with one p-bit
    for both endpoints) are fitted with error 0 at quality 0, and the packed
    block decodes to the input exactly: the axis, k-means, least squares,
    grid clamp and Pack all have to land on the two colours."""
    rng = np.random.default_rng(40 + mode)
    nsub = SPEC[mode][5]
    pbt = SPEC[mode][2]
    for trial in range(60):
        shape = int(rng.integers(16 if mode == 0 else 64)) if nsub > 1 else 0
        px = np.zeros(16, np.uint32)
        for s in range(nsub):
            pbit = int(rng.integers(2)) if pbt else -1
            c0 = _grid_colour(rng, mode, pbit)
            c1 = _grid_colour(rng, mode, pbit)
            mem = _members(tables, nsub, shape, s)
            pick = rng.integers(0, 2, len(mem))
            if len(mem) >= 2:
                pick[0], pick[1] = 0, 1
            for i, k in zip(mem, pick):
                px[i] = c1 if k else c0
        out, err = oracle.bc7_encode_candidate(px, mode, shape)
        assert err == 0.0, (mode, shape, px)
        assert np.array_equal(oracle.bc7_decode_block_spec(out), px), (mode, shape)


# --------------------------------------------------------------------------
# The rotation fit (modes 4 and 5)
# --------------------------------------------------------------------------

def test_rotation_fit_stale_statistics(oracle):
    """The rotated copy keeps the unrotated cluster's average, minimum and
    maximum (Compressor.cpp:664-695).  Mode 5, rotation 0: red alternates
    100/102, green 50, blue 60, alpha 10..25.  The stale average keeps alpha
    at 17.5 while every rotated point has alpha 255, so the covariance about
    it has an alpha-alpha term of 16 * 237.5^2 / 3 against 16/3 for red; the
    axis is alpha, all points project to one value, k-means fills one bucket
    and the colour becomes the single-colour fit of (101, 50, 60, 255): the
    block decodes to that flat colour where a fresh average would have kept
    the red pattern."""
    px = np.array([_pix(100 + 2 * (i % 2), 50, 60, 10 + i) for i in range(16)], np.uint32)
    e, p1, p2, idx, aidx, tr = oracle.fit_alpha(px, 5, 0)
    assert tr.path == oracle.FIT_ONE_BUCKET
    assert np.allclose(tr.axis, [0, 0, 0, 1], atol=1e-6)
    assert np.array_equal(np.array(tr.single), [101, 50, 60, 255])
    assert (idx == 1).all()
    out, _ = oracle.bc7_encode_candidate(px, 5, 0)
    dec = oracle.bc7_decode_block_spec(out)
    assert ((dec & 0xFFFFFF) == _pix(101, 50, 60, 0)).all()
    # with alpha 255 already (nothing stale), the rotated fit is the plain
    # RGB fit of the same points: the red pattern survives
    px2 = px | np.uint32(0xFF000000)
    e2, q1, q2, idx2, aidx2, tr2 = oracle.fit_alpha(px2, 5, 0)
    assert tr2.path == oracle.FIT_LEAST_SQUARES
    out2, _ = oracle.bc7_encode_candidate(px2, 5, 0)
    assert np.array_equal(oracle.bc7_decode_block_spec(out2), px2)


# Hand-derived from the reference's literal Optimal6CompressDXT1 rows
# (BCLookupTables.h:710):
#   [128] = {{0, 16, 63}, {1, 0, 63}}:  [0][0] == 0 -> a1 = 16<<2 | 16>>4 = 65,
#          a2 = 63<<2 | 16>>4 = 253 (the low bits come from [0][1], not [0][2])
#   [200] = {{0, 48, 52}, {1, 36, 63}}: a1 = 48<<2 | 3 = 195, a2 = 52<<2 | 3 = 211
#   [253] = {{1, 62, 63}, {0, 62, 63}}: [0][0] != 0 -> a1 = [1][1]<<2 | [0][1]>>4
#          = 62<<2 | 3 = 251, a2 = [1][2]<<2 | [0][1]>>4 = 63<<2 | 3 = 255
# Index 2 of 3-bit alpha indices (index mode 0) weighs 18/64, index 1 of
# 2-bit ones (index mode 1) 21/64; the error is 16 (v - interp)^2.
#   v=128: (65*46 + 253*18 + 32) >> 6 = 118 -> 16*100 = 1600;
#          (65*43 + 253*21 + 32) >> 6 = 127 -> 16
#   v=200: 200 and 200 -> 0, 0
#   v=253: (251*46 + 255*18 + 32) >> 6 = 252 -> 16; (251*43 + 255*21 + 32) >> 6 = 252 -> 16
EQUAL_ALPHA = {(128, 0): (65, 253, 2, 1600), (128, 1): (65, 253, 1, 16),
               (200, 0): (195, 211, 2, 0), (200, 1): (195, 211, 1, 0),
               (253, 0): (251, 255, 2, 16), (253, 1): (251, 255, 1, 16)}


@pytest.mark.parametrize("v,im", sorted(EQUAL_ALPHA))
def test_rotation_fit_equal_alpha_mode4(oracle, v, im):
    """Mode 4 with one alpha value takes its endpoints from the 6-bit DXT1
    table with the reference's mixed entries (Compressor.cpp:717-747).

    The colours are black and white halves; the rotated RGB fit lands on
    (0,0,0,255) and (255,255,255,255).  Its error is a closed form with two
    more reference behaviours in it:
      * mode 4 has no p-bits, yet QuantizedError quantises with the p-bit
        pair {0, 0} (CompressionMode.h:243-250, RGBAEndpoints.cpp:207-209):
        5-bit colour + 0 is a 6-bit grid with a zero lowest bit (255 -> 251)
        and 6-bit alpha + 0 a 7-bit one (255 -> 253);
      * the error compares against the cluster's stored pixels
        (GetPixel, RGBAEndpoints.cpp:231,271), which the rotated copy never
        updates, so the original alpha v counts against 253.
    So the RGB part is 8 * 3 * 4^2 + 16 * (v - 253)^2, plus the alpha term."""
    a1, a2, ai, aerr = EQUAL_ALPHA[(v, im)]
    px = np.array([_pix(*((255, 255, 255) if i % 2 else (0, 0, 0)), v) for i in range(16)],
                  np.uint32)
    e, p1, p2, idx, aidx, tr = oracle.fit_alpha(px, 4, 0, im)
    assert np.array_equal(tr.grid1, [0, 0, 0, 255]) and np.array_equal(tr.grid2, [255] * 4)
    assert (p1[3], p2[3]) == (a1, a2)
    assert (aidx == ai).all()
    assert e == 8 * 3 * 16 + 16 * (v - 253) ** 2 + aerr


@pytest.mark.parametrize("mode,rot,im", [(5, 0, 0), (5, 2, 0), (4, 0, 0), (4, 0, 1),
                                         (4, 3, 1)])
def test_rotation_fit_two_level_alpha_exact(oracle, mode, rot, im):
    """The alpha k-means and least squares (Compressor.cpp:760-905): a
    rotated channel with two values on the mode's alpha grid is fitted
    exactly, whatever the split."""
    rng = np.random.default_rng(mode * 10 + rot + im)
    ab = SPEC[mode][1]
    for _ in range(30):
        lo, hi = sorted(rng.choice(1 << ab, 2, replace=False))
        av = [_expand(int(lo), ab, -1), _expand(int(hi), ab, -1)]
        pick = rng.integers(0, 2, 16)
        pick[:2] = [0, 1]
        ch = np.zeros((16, 4), int)
        ch[:, 3] = 255
        ch[:, :3] = 255 if mode == 4 else 0
        ch[::2, :3] = 0
        target = 3 if rot == 0 else rot - 1
        ch[:, target] = [av[k] for k in pick]
        if rot:
            ch[:, 3] = 255  # the rotated-in channel carries the two levels
        px = (ch[:, 0] | ch[:, 1] << 8 | ch[:, 2] << 16 | ch[:, 3] << 24).astype(np.uint32)
        e, p1, p2, idx, aidx, tr = oracle.fit_alpha(px, mode, rot, im)
        # the least-squares levels, to float32 rounding
        assert np.allclose(sorted((p1[3], p2[3])), av, atol=1e-3), (p1[3], p2[3], av)
        # packed (endpoints rounded by ToPixel), the rotated channel decodes
        # to the input levels exactly
        out, _ = oracle.bc7_encode_candidate(px, mode, rot * 2 + im if mode == 4 else rot)
        dec = oracle.bc7_decode_block_spec(out)
        got = (dec >> np.uint32(8 * target)) & np.uint32(0xFF)
        assert np.array_equal(got, ch[:, target]), (mode, rot, im)


def test_rotation_fit_alpha_stages(oracle):
    """The alpha half of the rotation fit (Compressor.cpp:768-916) on random
    and generator blocks, every rotation and index mode: the 1-D k-means
    ends at a fixed point (first strictly nearest level, levels the exact
    bucket means, empty ones 0), the endpoints are the float64 least squares
    of its buckets clamped to [0, 255], they are quantised after truncation
    to uint8 (not rounding, as Pack later does), and the alpha error is the
    sum of each value's smallest squared distance to an interpolated level."""
    rng = np.random.default_rng(12)
    blocks = _random_blocks(rng, 120)
    checked = 0
    for b in blocks:
        ch = np.array([_channels(p) for p in b])
        for mode in (4, 5):
            ab = SPEC[mode][1]
            amask = (0xFF << (8 - ab)) & 0xFF
            for rot in range(4):
                for im in ((0, 1) if mode == 4 else (0,)):
                    e, p1, p2, idx, aidx, tr = oracle.fit_alpha(b, mode, rot, im)
                    if tr.alpha_path != oracle.FIT_LEAST_SQUARES:
                        continue
                    if (ch == ch[0]).all():
                        continue
                    av = ch[:, 3 if rot == 0 else rot - 1]
                    abits = SPEC[mode][4] if im == 0 else SPEC[mode][3]
                    nb = 1 << abits
                    cent = np.array(tr.alpha_centroid, np.float64)[:nb]
                    count = np.array(tr.alpha_count)[:nb]
                    bucket = np.array(tr.alpha_bucket)
                    assert np.array_equal(np.bincount(bucket, minlength=nb), count)
                    for j in range(nb):
                        if count[j]:
                            want = av[bucket == j].sum().astype(np.float32) / np.float32(count[j])
                            assert np.float32(cent[j]) == want
                        else:
                            assert cent[j] == 0
                    for i in range(16):
                        d = np.abs(av[i] - cent)
                        order = np.argsort(d, kind="stable")
                        if d[order[1]] - d[order[0]] <= 1e-4 and d[order[1]] != d[order[0]]:
                            continue
                        assert bucket[i] == order[0]
                    a = (nb - 1 - np.arange(nb)) / (nb - 1)
                    w = np.arange(nb) / (nb - 1)
                    asq, bsq, abw = (count * a * a).sum(), (count * w * w).sum(), \
                        (count * a * w).sum()
                    ax, bx = (count * a) @ cent, (count * w) @ cent
                    det = asq * bsq - abw * abw
                    l1 = np.clip((ax * bsq - bx * abw) / det, 0, 255)
                    l2 = np.clip((bx * asq - ax * abw) / det, 0, 255)
                    scale = (abs(ax * bsq) + abs(bx * abw) + abs(bx * asq) + abs(ax * abw)) / det
                    tol = 2e-5 * scale + 1e-3
                    assert abs(tr.alpha_lsq[0] - l1) <= tol and abs(tr.alpha_lsq[1] - l2) <= tol
                    q = [oracle.lib().ora_quantize_channel(int(x), amask, -1) for x in tr.alpha_lsq]
                    assert list(tr.alpha_q) == q  # int(): truncation
                    wts = WEIGHTS[abits]
                    lv = np.array([((64 - wj) * q[0] + wj * q[1] + 32) >> 6 for wj in wts])
                    err = sum(float(((av[i] - lv) ** 2).min()) for i in range(16))
                    assert tr.alpha_error == err
                    assert e == tr.rgb_error + tr.alpha_error
                    assert np.array_equal(aidx, [int(np.argmin((av[i] - lv) ** 2))
                                                 for i in range(16)])
                    checked += 1
    assert checked > 600, checked


# --------------------------------------------------------------------------
# BoxSelection
# --------------------------------------------------------------------------

def _estimate(oracle, px, members, nb):
    """EstimateTwo/ThreeClusterError (Compressor.cpp:1626-1655) from the
    definition: 0 for a degenerate bounding box, else 0.0001 plus the
    QuantizedError of the box corners."""
    pts = np.array([_channels(px[i]) for i in members])
    lo, hi = pts.min(axis=0), pts.max(axis=0)
    if ((hi - lo) ** 2).sum() == 0:
        return 0.0
    sel = np.array(members, np.uint8)
    mn = lo.astype(np.float32)
    mx = hi.astype(np.float32)
    return 0.0001 + oracle.lib().ora_quantized_error(px.ctypes.data, sel.ctypes.data,
                                                     len(members), mn.ctypes.data,
                                                     mx.ctypes.data, nb, 0xFFFFFFFF, None,
                                                     None)


def _box_selection(oracle, tables, px):
    opaque = all((int(p) >> 24) >= 250 for p in px)
    best2 = best3 = None
    sh2 = sh3 = 0
    for i in range(64):
        err = 0.0
        for s in range(2):
            err += _estimate(oracle, px, _members(tables, 2, i, s), 8)
        if best2 is None or err < best2:
            best2, sh2 = err, i
        if err < 1e-9:
            return 1, sh2, None, 0b10001010
    if not opaque:
        return 1, sh2, None, 0b11110000
    for i in range(64):
        err = 0.0
        for s in range(3):
            err += _estimate(oracle, px, _members(tables, 3, i, s), 4)
        if best3 is None or err < best3:
            best3, sh3 = err, i
        if err < 1e-9:
            return 2, sh2, sh3, 0b00000101
    return 2, sh2, sh3, 0b11001111


def test_box_selection_rescan(oracle, tables):
    """BoxSelection (Compressor.cpp:1670-1750): the chosen two- and
    three-subset shapes, shape count and modes equal a rescan of all 64
    shapes of each kind from the definitions (first strict minimum,
    early exit below 1e-9)."""
    rng = np.random.default_rng(5)
    blocks = _random_blocks(rng, 120)
    s2 = tables[0]
    s3 = tables[1]
    for k in range(40):  # two flat colours on a random two-subset shape
        sh = int(rng.integers(64))
        c = rng.integers(0, 2**24, 2) | 0xFF000000
        blocks.append(np.array([c[(s2[sh] >> i) & 1] for i in range(16)], np.uint32))
    for k in range(40):  # three flat colours on a three-subset shape
        sh = int(rng.integers(64))
        c = rng.integers(0, 2**24, 3) | 0xFF000000
        blocks.append(np.array([c[_subset_of(tables, 3, sh, i)] for i in range(16)],
                               np.uint32))
    for k in range(60):  # a few flat colours scattered: ties between shapes
        c = (rng.integers(0, 2**24, 3) | 0xFF000000).astype(np.uint32)
        c[1] = c[0] ^ np.uint32(rng.integers(1, 4))
        blocks.append(c[rng.integers(0, 3, 16)].astype(np.uint32))
    for k in range(30):  # alpha just opaque (250..254) and just not (249)
        b = (rng.integers(0, 2**24, 16) | (rng.integers(250, 256, 16) << 24)).astype(np.uint32)
        if k % 3 == 0:
            b[int(rng.integers(16))] = (b[0] & 0xFFFFFF) | (249 << 24)
        blocks.append(b)
    seen3 = set()
    early = {2: 0, 3: 0}
    for b in blocks:
        ns, sh2, sh3, modes = _box_selection(oracle, tables, b)
        n, g2, g3, gm = (ctypes.c_uint32() for _ in range(4))
        oracle.lib().ora_bc7_box_selection(b.ctypes.data, ctypes.byref(n), ctypes.byref(g2),
                                           ctypes.byref(g3), ctypes.byref(gm))
        assert (n.value, g2.value, gm.value) == (ns, sh2, modes), (b, ns, sh2, modes)
        if sh3 is not None:
            assert g3.value == sh3
            seen3.add(sh3)
        if modes == 0b10001010:
            early[2] += 1
        if modes == 0b00000101:
            early[3] += 1
    assert early[2] >= 30 and early[3] >= 20 and len(seen3) >= 20, (early, len(seen3))


# --------------------------------------------------------------------------
# Pack and the specification decoder
# --------------------------------------------------------------------------

class _Bits:
    def __init__(self, blk):
        self.v = int.from_bytes(bytes(blk), "little")
        self.pos = 0

    def take(self, n):
        r = (self.v >> self.pos) & ((1 << n) - 1)
        self.pos += n
        return r


def spec_decode(tables, blk):
    """A BC7 block decoder written from the format specification
    (independent of the oracle): returns 16 RGBA8 pixels and the fields."""
    br = _Bits(blk)
    mode = 0
    while mode < 8 and br.take(1) == 0:
        mode += 1
    if mode == 8:
        return np.zeros(16, np.uint32), None
    cb, ab, pbt, ib, aib, nsub, pb = SPEC[mode]
    shape = br.take(pb)
    rot = br.take(2) if mode in (4, 5) else 0
    im = br.take(1) if mode == 4 else 0
    ep = np.zeros((nsub, 2, 4), int)
    for ch in range(3):
        for s in range(nsub):
            for e in range(2):
                ep[s, e, ch] = br.take(cb)
    for s in range(nsub):
        for e in range(2):
            ep[s, e, 3] = br.take(ab) if ab else 0
    pbits = np.full((nsub, 2), -1)
    if pbt == 2:
        for s in range(nsub):
            for e in range(2):
                pbits[s, e] = br.take(1)
    elif pbt == 1:
        for s in range(nsub):
            pbits[s, :] = br.take(1)
    full = np.zeros((nsub, 2, 4), int)
    for s in range(nsub):
        for e in range(2):
            for ch in range(4):
                bits = cb if ch < 3 else ab
                if bits == 0:
                    full[s, e, ch] = 255
                else:
                    full[s, e, ch] = _expand(ep[s, e, ch], bits, int(pbits[s, e]))
    anchors = {_anchor(tables, nsub, shape, s) for s in range(nsub)}
    first = [br.take(ib - 1 if i in anchors else ib) for i in range(16)]
    second = [br.take(aib - 1 if i == 0 else aib) for i in range(16)] if aib else None
    if mode == 4 and im == 1:
        cidx, aidx, cbits, abits = second, first, aib, ib
    elif aib:
        cidx, aidx, cbits, abits = first, second, ib, aib
    else:
        cidx, aidx, cbits, abits = first, first, ib, ib
    out = np.zeros(16, np.uint32)
    for i in range(16):
        s = _subset_of(tables, nsub, shape, i)
        c = []
        for ch in range(4):
            w = WEIGHTS[abits][aidx[i]] if ch == 3 else WEIGHTS[cbits][cidx[i]]
            e0, e1 = full[s, 0, ch], full[s, 1, ch]
            c.append(((64 - w) * e0 + w * e1 + 32) >> 6)
        if rot:
            c[rot - 1], c[3] = c[3], c[rot - 1]
        out[i] = _pix(*c)
    return out, dict(mode=mode, shape=shape, rot=rot, im=im)


def _interp_pixels(tables, mode, shape, rot, im, ends, idx, aidx):
    """Pixels of BC7 parameters by the specification's interpolation."""
    cb, ab, pbt, ib, aib, nsub, pb = SPEC[mode]
    cbits, abits = (aib, ib) if (mode == 4 and im == 1) else (ib, aib if aib else ib)
    out = np.zeros(16, np.uint32)
    for i in range(16):
        s = _subset_of(tables, nsub, shape, i)
        c = []
        for ch in range(4):
            w = WEIGHTS[abits][aidx[i] if aib else idx[i]] if ch == 3 else WEIGHTS[cbits][idx[i]]
            e0, e1 = ends[s][0][ch], ends[s][1][ch]
            c.append(((64 - w) * e0 + w * e1 + 32) >> 6)
        if rot:
            c[rot - 1], c[3] = c[3], c[rot - 1]
        out[i] = _pix(*c)
    return out


@pytest.mark.parametrize("mode", range(8))
def test_pack_round_trip(oracle, tables, mode):
    """Pack (Compressor.cpp:1096-1298) of random parameters on the mode's grid
    decodes, by the specification reader, to the parameters' pixels.  Pack
    may swap a subset's endpoints and invert its indices (anchor bit); the
    pixels stay the same, except that the reference leaves the p-bits
    unswapped (Compressor.cpp:1149-1170), so unequal p-bits are only drawn
    where no swap happens."""
    rng = np.random.default_rng(60 + mode)
    cb, ab, pbt, ib, aib, nsub, pb = SPEC[mode]
    combos = _combos(pbt)
    for trial in range(200):
        shape = int(rng.integers(1 << pb)) if pb else 0
        rot = int(rng.integers(4)) if mode in (4, 5) else 0
        im = int(rng.integers(2)) if mode == 4 else 0
        cbits, abits = (aib, ib) if (mode == 4 and im == 1) else (ib, aib)
        idx = rng.integers(0, 1 << cbits, 16).astype(np.uint8)
        aidx = rng.integers(0, 1 << abits, 16).astype(np.uint8) if aib else np.zeros(16, np.uint8)
        no_swap = trial % 2 == 0
        if no_swap:
            for s in range(nsub):
                idx[_anchor(tables, nsub, shape, s)] &= (1 << (cbits - 1)) - 1
            if aib:
                aidx[0] &= (1 << (abits - 1)) - 1
        ep = np.zeros((3, 2, 4), np.float32)
        ends = []
        pbc = np.zeros(3, np.uint8)
        for s in range(nsub):
            if pbt == 2 and not no_swap:
                ci = int(rng.choice([0, 3]))  # equal p-bits
            else:
                ci = int(rng.integers(len(combos)))
            pbc[s] = ci
            pp = combos[ci]
            pair = []
            for e in range(2):
                vals = []
                for ch in range(4):
                    bits = cb if ch < 3 else ab
                    if bits == 0:
                        vals.append(255)
                    else:
                        vals.append(_expand(int(rng.integers(1 << bits)), bits, pp[e]))
                ep[s, e] = vals
                pair.append(vals)
            ends.append(pair)
        out = oracle.bc7_pack(mode, shape, rot, im, ep, pbc, idx, aidx)
        dec, f = spec_decode(tables, out)
        assert f == dict(mode=mode, shape=shape, rot=rot, im=im)
        want = _interp_pixels(tables, mode, shape, rot, im, ends, idx, aidx)
        assert np.array_equal(dec, want), (mode, trial)
        assert np.array_equal(oracle.bc7_decode_block_spec(out), dec)


def test_spec_decoder_agrees_on_random_blocks(oracle, tables):
    """The oracle's specification decoder equals the independent reader on
    random blocks of every mode (and the reserved all-zero mode byte)."""
    rng = np.random.default_rng(9)
    for k in range(2000):
        blk = rng.integers(0, 256, 16).astype(np.uint8)
        mode = k % 9
        v = int.from_bytes(blk.tobytes(), "little")
        v = (v >> (mode + 1) << (mode + 1)) | ((1 << mode) if mode < 8 else 0)
        if mode == 8:
            v &= ~0xFF
        blk = np.frombuffer(v.to_bytes(16, "little"), np.uint8).copy()
        want, _ = spec_decode(tables, blk)
        assert np.array_equal(oracle.bc7_decode_block_spec(blk), want), (mode, blk)
