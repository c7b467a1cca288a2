"""BC7 parity: HIP encoder vs the CPU oracle on generated images.

The bar is bit-exact output: at quality 0 the search is deterministic; at
quality > 0 both sides use the same counter-based annealing stream.  At the
full 4096x4096 size, where the oracle is too slow, size-independent
properties are checked instead (sampled blocks against the oracle, no block
worse than quality 0, PSNR, sub-job composition).
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


def _wm():
    return ctypes.c_uint32(0xFFFFFFFF)


@pytest.mark.parametrize("w,h,seed", [(64, 64, 1), (128, 128, 7), (256, 256, 1), (96, 32, 4),
                                      (4, 4, 2), (4, 64, 8)])
def test_bc7_q0_bit_exact(fh, oracle, w, h, seed):
    """Quality 0 (the 256x256 seed-1 case is the reference's CPU config)."""
    img = synth.generate(w, h, seed)
    print(f"bc7 q0 {w}x{h} seed={seed}", flush=True)
    ref, ref_wm = oracle.bc7_compress(img, quality=0)
    wm = _wm()
    got = fh.compress(img, quality=0, wm_counter=wm)
    bad = _first_diff(got, ref)
    assert bad.size == 0, f"{bad.size} blocks differ, first {bad[:8]}"
    assert wm.value == ref_wm


@pytest.mark.parametrize("guard", [False, True])
def test_bc7_q50_counter_rng_bit_exact(fh, oracle, guard):
    """Annealing with the shared counter stream, with and without the
    quality guard."""
    img = synth.generate(64, 64, 3)
    print(f"bc7 q50 64x64 seed=3 guard={guard}", flush=True)
    ref, _ = oracle.bc7_compress(img, quality=50, seed=1234, q0_guard=guard)
    got = fh.compress(img, quality=50, seed=1234, wm_counter=_wm(), q0_guard=guard)
    bad = _first_diff(got, ref)
    assert bad.size == 0, f"{bad.size} blocks differ, first {bad[:8]}"


@pytest.mark.parametrize("q", [1, 2, 7, 256])
def test_bc7_other_qualities_bit_exact(fh, oracle, q):
    """Annealing edge cases: q=1 (temperature 0/0), q=2, odd q, q=256."""
    img = synth.generate(32, 32, 12)
    print(f"bc7 q={q} 32x32", flush=True)
    ref, _ = oracle.bc7_compress(img, quality=q, seed=99, q0_guard=True)
    got = fh.compress(img, quality=q, seed=99, wm_counter=_wm())
    bad = _first_diff(got, ref)
    assert bad.size == 0, f"{bad.size} blocks differ, first {bad[:8]}"


@pytest.mark.parametrize("modes", [0x01, 0x40, 0x30, 0x8A, 0x7F])
def test_bc7_block_mode_masks(fh, oracle, modes):
    """BPTCC::CompressionSettings::m_BlockModes restricts the search."""
    img = synth.generate(32, 32, 5)
    print(f"bc7 modes=0x{modes:02x}", flush=True)
    ref, _ = oracle.bc7_compress(img, quality=4, seed=3, block_modes=modes, q0_guard=True)
    got = fh.compress(img, quality=4, seed=3, block_modes=modes, wm_counter=_wm())
    bad = _first_diff(got, ref)
    assert bad.size == 0, f"{bad.size} blocks differ, first {bad[:8]}"


def test_bc7_sub_jobs_compose(fh, oracle):
    """Sub-rectangle jobs (CompressionJob XStart..YEnd) in raster order give
    the whole-image bytes, the watermark counter carrying across jobs."""
    img = synth.generate(64, 48, 21)
    whole = fh.compress(img, quality=0, wm_counter=_wm())
    out = np.zeros_like(whole)
    wm = _wm()
    for rect in [(0, 0, 64, 8), (0, 12, 32, 24), (32, 24, 64, 48)]:
        fh.compress(img, quality=0, rect=rect, out=out, wm_counter=wm)
    assert np.array_equal(out, whole)
    ref, _ = oracle.bc7_compress(img, quality=0, rect=(0, 12, 32, 24))
    blocks = ref.reshape(-1, 16)
    got = fh.compress(img, quality=0, rect=(0, 12, 32, 24), wm_counter=_wm()).reshape(-1, 16)
    assert np.array_equal(got[48:104], blocks[48:104])


def test_bc7_solid_and_transparent_images(fh, oracle):
    img = np.full((16, 16), 0xFF336699, dtype=np.uint32)
    img[8:, :] = 0x00123456  # transparent bottom half
    ref, rwm = oracle.bc7_compress(img, quality=50, seed=1)
    wm = _wm()
    got = fh.compress(img, quality=50, seed=1, wm_counter=wm)
    assert np.array_equal(got, ref) and wm.value == rwm


def test_bc7_decode_matches_oracle(fh, oracle):
    img = synth.generate(64, 64, 5)
    data, _ = oracle.bc7_compress(img, quality=0)
    dec_ref = oracle.bc7_decompress(data, 64, 64)
    dec = fh.decompress(data, 64, 64)
    assert np.array_equal(dec, dec_ref)


def _block_premul_err(img, dec):
    """Per-block sum of (a*c - a'*c')^2 over RGB (Image::ComputePSNR's term
    scaled by 255^2), int64, in raster block order."""
    h, w = img.shape
    a = img.view(np.uint8).reshape(h, w, 4).astype(np.int64)
    d = dec.view(np.uint8).reshape(h, w, 4).astype(np.int64)
    e = ((a[..., :3] * a[..., 3:4] - d[..., :3] * d[..., 3:4]) ** 2).sum(-1)
    return e.reshape(h // 4, 4, w // 4, 4).sum((1, 3)).reshape(-1)


def test_bc7_4096_q50_properties(fh, oracle):
    """The benchmark workload (4096x4096 seed 3, q=50): two half-image
    jobs compose, sampled blocks match the oracle, no block decodes worse
    than its quality-0 encoding, and annealing does not lower PSNR."""
    n = 4096
    img = synth.generate(n, n, 3)
    print("bc7 4096 q50", flush=True)
    wm = _wm()
    q50 = fh.compress(img, quality=50, seed=3, wm_counter=wm)
    q0 = fh.compress(img, quality=0, wm_counter=_wm())
    d50 = fh.decompress(q50, n, n)
    d0 = fh.decompress(q0, n, n)
    e50 = _block_premul_err(img, d50)
    e0 = _block_premul_err(img, d0)
    assert int((e50 > e0).sum()) == 0
    p50, p0 = oracle.psnr(img, d50), oracle.psnr(img, d0)
    print(f"psnr q0 {p0:.4f} q50 {p50:.4f}", flush=True)
    assert p50 >= p0
    # halves as separate jobs: same bytes
    out = np.zeros_like(q50)
    wm2 = _wm()
    # rows up to and including y_end are encoded, so the first half ends at
    # the last block row before n/2
    fh.compress(img, quality=50, seed=3, rect=(0, 0, n, n // 2 - 4), out=out, wm_counter=wm2)
    fh.compress(img, quality=50, seed=3, rect=(0, n // 2, n, n), out=out, wm_counter=wm2)
    assert np.array_equal(out, q50) and wm2.value == wm.value
    # sampled blocks against the oracle, solid blocks with their watermark
    blocks = synth.blocks(img)
    solid = (blocks == blocks[:, :1]).all(axis=1)
    nsolid_before = np.concatenate([[0], np.cumsum(solid)[:-1]])
    rng = np.random.default_rng(0)
    picks = np.concatenate([rng.integers(0, blocks.shape[0], 300),
                            [0, (1 << 19) - 1, 1 << 19, blocks.shape[0] - 1]])
    got = q50.reshape(-1, 16)
    for b in picks:
        wmi = int((nsolid_before[b]) % 9)
        ref, _ = oracle.bc7_compress_block(blocks[b], quality=50, seed=3, block_id=int(b),
                                           wm_index=wmi, q0_guard=True)
        assert np.array_equal(got[b], ref), int(b)


def test_bc7_multi_call_job_q0(fh, oracle):
    """A job above the per-call block limit (4096x8192 = 2^21 blocks) runs as
    consecutive calls; sampled blocks around the call boundary match the
    oracle and the watermark counter ends where the reference's would."""
    w, h = 4096, 8192
    img = synth.generate(w, h, 6)
    print("bc7 q0 4096x8192", flush=True)
    wm = _wm()
    got = fh.compress(img, quality=0, wm_counter=wm).reshape(-1, 16)
    blocks = synth.blocks(img)
    solid = (blocks == blocks[:, :1]).all(axis=1)
    nsolid_before = np.concatenate([[0], np.cumsum(solid)[:-1]])
    assert wm.value == (int(solid.sum()) - 1) % 9
    rng = np.random.default_rng(1)
    picks = np.concatenate([rng.integers(0, blocks.shape[0], 200),
                            np.arange((1 << 20) - 40, (1 << 20) + 40)])
    for b in picks:
        ref, _ = oracle.bc7_compress_block(blocks[b], quality=0, block_id=int(b),
                                           wm_index=int(nsolid_before[b] % 9))
        assert np.array_equal(got[b], ref), int(b)


def test_bc7_q50_psnr_vs_reference_statistics(fh, oracle):
    """PSNR at q=50 without the quality guard (the reference algorithm) is
    within 0.05 dB of the oracle run with the reference's own LCG annealing
    stream, or better: mean of five GPU seeds against the mean of five
    time-seed-like LCG seeds, both without the guard (it is tested on its
    own).  The reference seeds its LCG with time(NULL) (Compressor.cpp:358),
    so only the distribution can be compared."""
    from concurrent.futures import ThreadPoolExecutor
    img = synth.generate(256, 256, 1)
    gpu = []
    for s in (11, 12, 13, 14, 15):
        got = fh.compress(img, quality=50, seed=s, wm_counter=_wm(), q0_guard=False)
        gpu.append(oracle.psnr(img, fh.decompress(got, 256, 256)))

    def ref_run(s):
        ref, _ = oracle.bc7_compress(img, quality=50, rng=oracle.RNG_LCG, lcg_seed=s)
        return oracle.psnr(img, oracle.bc7_decompress(ref, 256, 256))

    with ThreadPoolExecutor(5) as ex:
        refs = list(ex.map(ref_run, (1, 2, 3, 1764000000, 1764000001)))
    print(f"gpu {np.mean(gpu):.4f} +- {np.std(gpu):.4f}, ref {np.mean(refs):.4f} +- "
          f"{np.std(refs):.4f}", flush=True)
    assert np.mean(gpu) >= np.mean(refs) - 0.05
