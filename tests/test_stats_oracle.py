"""CompressWithStats on the CPU: the oracle's statistics block path
(BPTCEncoder/src/Compressor.cpp:1954-2183) and the log text of
fh_bc7_stats_format (BlockLogger and PrintStat, Compressor.cpp:1373-1390,
1945-1950), without a GPU.

The oracle's scan is pinned by known answers (solid, transparent and
two-colour blocks whose estimates follow from the definitions), by invariants
the reference's code fixes for every block (mode errors stay -1, logged
estimates strictly decrease, the opaque test is A == 255), and by a brute-force
rescan built on the pinned QuantizedError (ora_quantized_error, checked in
tests/test_oracle_bc7.py).  The text format is pinned by hand-written lines in
the form the reference's ostream writes them."""
import ctypes

import numpy as np
import pytest

import fastc_amd
from fastc_amd import synth

NAMES = ["Zero", "One", "Two", "Three", "Four", "Five", "Six", "Seven"]


def _mask_bits(shape, nparts, oracle):
    s2 = np.zeros(64, np.uint16)
    s3 = np.zeros(128, np.uint16)
    a2 = np.zeros(64, np.uint8)
    a3 = np.zeros(128, np.uint8)
    iv = np.zeros(128, np.uint32)
    oracle.lib().ora_table_partitions(s2.ctypes.data, s3.ctypes.data, a2.ctypes.data,
                                      a3.ctypes.data, iv.ctypes.data)
    if nparts == 2:
        m = int(s2[shape])
        return [[i for i in range(16) if ((m >> i) & 1) == s] for s in range(2)]
    m1, m2 = int(s3[2 * shape]), int(s3[2 * shape + 1])
    sub = [((m1 >> i) & 1) + ((m2 >> i) & 1) if (m1 >> i) & 1 else 0 for i in range(16)]
    return [[i for i in range(16) if sub[i] == s] for s in range(3)]


def _qerr(oracle, px, sel, nb, mask):
    """QuantizedError of the points `sel` between their bounding-box corners."""
    pts = np.array([[(int(px[i]) >> (8 * k)) & 255 for k in range(4)] for i in sel], np.float32)
    lo = pts.min(0).astype(np.float32)
    hi = pts.max(0).astype(np.float32)
    if (lo == hi).all():
        return None
    s = np.array(sel, np.uint8)
    return oracle.lib().ora_quantized_error(px.ctypes.data, s.ctypes.data, len(sel),
                                            lo.ctypes.data, hi.ctypes.data, nb, mask, None, None)


def _rescan(oracle, px):
    """The statistics scan restated from its definition, per shape and subset
    with ora_quantized_error: (logged two, logged three, path, estimates)."""
    est = [-1.0] * 8
    opaque = all((int(p) >> 24) == 255 for p in px)
    allpts = list(range(16))
    e6 = _qerr(oracle, px, allpts, 4, 0xFEFEFEFE)
    est[6] = 0.0 if e6 is None else 0.0001 + e6
    logged = {2: [], 3: []}
    path = 3
    for nparts, nb, masks, modes in ((2, 8, (0xFFFCFCFC, 0xFFFEFEFE), (1, 3)),
                                     (3, 4, (0xFFF0F0F0, 0xFFF8F8F8), (0, 2))):
        if nparts == 3 and not opaque:
            break
        best = float("inf")
        for shape in range(64):
            err = 0.0
            sums = [None, None]
            for sel in _mask_bits(shape, nparts, oracle):
                es = [_qerr(oracle, px, sel, nb, m) for m in masks]
                if es[0] is None:
                    vals, x = [0.0, 0.0], 0.0
                else:
                    vals = [e if nparts == 2 else 0.0001 + e for e in es]
                    x = 0.0001
                    x += vals[1] if vals[1] < vals[0] else vals[0]
                err += x
                sums = [v if s is None else s + v for s, v in zip(sums, vals)]
            for m, v in zip(modes, sums):
                if est[m] == -1.0 or v < est[m]:
                    est[m] = v
            if err < best:
                logged[nparts].append(err)
                best = err
            if err < 1e-9:
                path = 2
                break
    return logged[2], logged[3], path, est


def test_solid_and_transparent_known_answers(oracle):
    """The solid path logs Path 0 and mode 5, the transparent path Path 1 and
    mode 6, before any estimate is made (Compressor.cpp:2019-2044)."""
    img = np.zeros((4, 8), np.uint32)
    img[:, :4] = 0xFF336699
    img[:, 4:] = np.arange(16, dtype=np.uint32).reshape(4, 4) * 0x010101  # alpha 0
    _, _, st = oracle.bc7_compress_logged(img)
    assert (st["path"][0], st["mode"][0]) == (0, 5)
    assert (st["path"][1], st["mode"][1]) == (1, 6)
    assert (st["num_two_shape"] == 0).all() and (st["num_three_shape"] == 0).all()
    assert (st["mode_estimate"] == -1).all() and (st["mode_error"] == -1).all()


def test_two_colour_block_stops_at_shape_zero(oracle):
    """Two flat colours laid out as shape 0 (mask 0xCCCC): both subsets are one
    colour, so shape 0's estimate is 0, the scan stops there (Path 2) and the
    estimates of modes 1 and 3 are 0; the opaque block still scans
    three-subset shapes."""
    m = 0xCCCC
    px = np.array([0xFF80A0C0 if (m >> i) & 1 else 0xFF102030 for i in range(16)], np.uint32)
    _, _, st = oracle.bc7_compress_logged(px.reshape(4, 4))
    r = st[0]
    assert r["path"] == 2 and r["num_two_shape"] == 1 and r["two_shape"][0] == 0.0
    assert r["mode_estimate"][1] == 0.0 and r["mode_estimate"][3] == 0.0
    assert r["num_three_shape"] >= 1
    # the early exit leaves modes {1, 3, 7}, or {0, 2} if a three-subset shape
    # is exact too
    assert r["mode"] in (0, 1, 2, 3, 7)


def test_scan_matches_brute_force_rescan(oracle):
    """Every statistic of the oracle's scan equals the restatement above on
    generator blocks, A in 250..254 blocks (opaque for BoxSelection, not here)
    and random blocks."""
    rng = np.random.default_rng(11)
    img = synth.generate(32, 32, 7)
    blocks = [img[y:y + 4, x:x + 4].reshape(16).copy() for y in range(0, 32, 4)
              for x in range(0, 32, 4)]
    for _ in range(6):
        b = blocks[int(rng.integers(len(blocks)))].copy()
        b = (b & 0x00FFFFFF) | (rng.integers(250, 255, 16).astype(np.uint32) << 24)
        blocks.append(b)
    for _ in range(6):
        blocks.append(rng.integers(0, 2**32, 16, dtype=np.uint64).astype(np.uint32))
    for px in blocks:
        if (px == px[0]).all() or not (px >> 24).any():
            continue
        _, _, st = oracle.bc7_compress_logged(px.reshape(4, 4))
        r = st[0]
        two, three, path, est = _rescan(oracle, px)
        assert list(r["two_shape"][:r["num_two_shape"]]) == two
        assert list(r["three_shape"][:r["num_three_shape"]]) == three
        assert r["path"] == path
        assert list(r["mode_estimate"]) == est


def test_invariants_on_an_image(oracle):
    img = synth.generate(64, 64, 2)
    out, _, st = oracle.bc7_compress_logged(img)
    blocks = img.reshape(16, 4, 16, 4).transpose(0, 2, 1, 3).reshape(-1, 16)
    for k, r in enumerate(st):
        assert (r["mode_error"] == -1).all()
        assert r["mode"] == next(m for m in range(8) if (out[16 * k] >> m) & 1)
        two = r["two_shape"][:r["num_two_shape"]]
        three = r["three_shape"][:r["num_three_shape"]]
        assert (np.diff(two) < 0).all() and (np.diff(three) < 0).all()
        opaque = ((blocks[k] >> 24) == 255).all()
        if r["path"] >= 2:
            assert r["num_two_shape"] >= 1
            assert (r["num_three_shape"] >= 1) == bool(opaque)
            if not opaque:
                assert r["mode_estimate"][0] == -1 and r["mode_estimate"][2] == -1
            assert (r["path"] == 2) == bool((two[-1:] < 1e-9).any() or (three[-1:] < 1e-9).any())


def test_stats_path_encodes_like_compress_with_mode_6_only(oracle):
    """Both block paths share CompressClusters: with mode 6 alone, whose fit
    does not depend on the chosen shapes, the statistics path writes the
    same bytes as Compress for every block (both scans stop at a shape whose
    subsets are all one colour, so their early exits agree)."""
    img = synth.generate(64, 64, 3)
    for q in (0, 6):
        a, _ = oracle.bc7_compress(img, block_modes=0x40, quality=q, seed=4)
        b, _, _ = oracle.bc7_compress_logged(img, block_modes=0x40, quality=q, seed=4)
        assert np.array_equal(a, b), q


def _fmt(L, rec, idx):
    size = ctypes.c_size_t()
    assert L.fh_bc7_stats_format(rec.ctypes.data, idx, None, 0, ctypes.byref(size)) == 0
    buf = ctypes.create_string_buffer(size.value + 1)
    assert L.fh_bc7_stats_format(rec.ctypes.data, idx, buf, size.value + 1,
                                 ctypes.byref(size)) == 0
    return buf.value.decode()


def test_format_known_answer():
    """The exact lines of a solid block and of a searched block, as the
    reference's ostream writes them: integers in decimal, doubles in the
    default (%g, six significant digits) format."""
    L = fastc_amd.lib()
    rec = np.zeros(1, fastc_amd.BLOCK_STATS_DTYPE)
    rec["mode_estimate"] = -1.0
    rec["mode_error"] = -1.0
    rec["path"] = 0
    rec["mode"] = 5
    want = "7: BlockStat_Path -- 0\n7: BlockStat_Mode -- 5\n" + "".join(
        "7: BlockStat_Mode%sEstimate -- -1\n7: BlockStat_Mode%sError -- -1\n" % (n, n)
        for n in NAMES)
    assert _fmt(L, rec, 7) == want
    rec["path"] = 3
    rec["mode"] = 1
    rec["num_two_shape"] = 3
    rec["two_shape"][0, :3] = [1234567.0002, 735.0006, 0.0002]
    rec["num_three_shape"] = 1
    rec["three_shape"][0, 0] = 12.0003
    rec["mode_estimate"][0, 6] = 99999.5
    text = _fmt(L, rec, 4294967296).splitlines()
    assert text[:6] == ["4294967296: BlockStat_TwoShapeEstimate -- 1.23457e+06",
                        "4294967296: BlockStat_TwoShapeEstimate -- 735.001",
                        "4294967296: BlockStat_TwoShapeEstimate -- 0.0002",
                        "4294967296: BlockStat_ThreeShapeEstimate -- 12.0003",
                        "4294967296: BlockStat_Path -- 3",
                        "4294967296: BlockStat_Mode -- 1"]
    assert "4294967296: BlockStat_ModeSixEstimate -- 99999.5" in text


def test_format_matches_oracle_text(oracle):
    """fh_bc7_stats_format equals the oracle's formatting for every block of an
    image, and reports the size and a too-small buffer."""
    L = fastc_amd.lib()
    img = synth.generate(64, 32, 5)
    _, _, st = oracle.bc7_compress_logged(img)
    for k in range(st.size):
        rec = np.ascontiguousarray(st[k:k + 1]).view(fastc_amd.BLOCK_STATS_DTYPE)
        assert _fmt(L, rec, k + 100) == oracle.format_block_stats(st[k], k + 100)
    size = ctypes.c_size_t()
    small = ctypes.create_string_buffer(8)
    rec = np.ascontiguousarray(st[:1]).view(fastc_amd.BLOCK_STATS_DTYPE)
    assert L.fh_bc7_stats_format(rec.ctypes.data, 0, small, 8, ctypes.byref(size)) == 5
    assert size.value > 8
    bad = rec.copy()
    bad["num_two_shape"] = 65
    assert L.fh_bc7_stats_format(bad.ctypes.data, 0, None, 0, ctypes.byref(size)) == 1


def test_stats_entry_point_validation():
    """Formats other than BPTC are unsupported (the reference only collects
    BPTC statistics); a short statistics buffer is rejected before any device
    work."""
    img = np.zeros((8, 8), np.uint32)
    out = np.zeros(64, np.uint8)
    j = fastc_amd._job(0, img.ctypes.data, out.ctypes.data, out.nbytes, 8, 8, None, 0, 0, 0xFF,
                       None, True, None, 4, 0.6)
    st = np.zeros(4, fastc_amd.BLOCK_STATS_DTYPE)
    L = fastc_amd.lib()
    assert L.fh_compress_with_stats(ctypes.byref(j), st.ctypes.data, 4) == 6
    j.format = 3
    j.out_size = 4 * 16
    assert L.fh_compress_with_stats(ctypes.byref(j), st.ctypes.data, 3) == 5
    assert L.fh_compress_with_stats(ctypes.byref(j), None, 0) == 5


def test_stats_device_entry_point_checks():
    """The device entry point rejects a short or misaligned record buffer
    before any device work."""
    img = np.zeros((8, 8), np.uint32)
    out = np.zeros(64, np.uint8)
    j = fastc_amd._job(3, img.ctypes.data, out.ctypes.data, out.nbytes, 8, 8, None, 0, 0, 0xFF,
                       None, True, None, 4, 0.6)
    L = fastc_amd.lib()
    assert L.fh_compress_with_stats_device(ctypes.byref(j), ctypes.c_void_p(4096), 3, None) == 5
    assert L.fh_compress_with_stats_device(ctypes.byref(j), ctypes.c_void_p(4100), 4, None) == 1
