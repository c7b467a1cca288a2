"""The C++ FasTC-style API (include/fasthip/fastc.hpp), compiled as a FasTC
caller would use it: error behaviour without a GPU, and on the GPU the same
bytes as the C ABI and the reference fixtures."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from fastc_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def facade_bin(tmp_path_factory):
    import fastc_amd
    fastc_amd.lib()
    out = tmp_path_factory.mktemp("facade") / "facade_test"
    libdir = os.path.join(ROOT, "fastc_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "facade_test.cpp"),
                    "-L", libdir, "-lfasthip", f"-Wl,-rpath,{libdir}",
                    "-Wl,--allow-shlib-undefined", "-pthread", "-o", str(out)], check=True)
    return str(out)


def test_facade_error_behaviour(facade_bin):
    r = subprocess.run([facade_bin, "validate"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "TexComp -- " in r.stderr


@pytest.mark.gpu
def test_facade_matches_c_abi(fh, facade_bin, tmp_path):
    img = synth.generate(64, 64, 2, edges=True)
    raw = tmp_path / "in.raw"
    img.tofile(raw)
    r = subprocess.run([facade_bin, "encode", str(raw), "64", "64", str(tmp_path)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    g = np.load(os.path.join(ROOT, "tests", "golden", "dxt_etc.npz"))
    rd = lambda n: np.fromfile(tmp_path / n, dtype=np.uint8)
    assert np.array_equal(rd("dxt1.bin"), g["dxt1"])
    assert np.array_equal(rd("dxt5.bin"), g["dxt5"])
    assert np.array_equal(rd("etc1.bin"), g["etc1"])
    want = fh.compress(img, quality=0, wm_counter=ctypes.c_uint32(0xFFFFFFFF))
    bptc = rd("bptc.bin").reshape(-1, 16)
    # identical except the solid-block watermark entries, which depend on the
    # process-wide counter (first call here vs a fresh counter above)
    from oracle import pyoracle
    dec_a = pyoracle.bc7_decompress(rd("bptc.bin"), 64, 64)
    dec_b = pyoracle.bc7_decompress(want, 64, 64)
    assert np.array_equal(dec_a, dec_b)
    assert np.array_equal(pyoracle.bc7_decompress(rd("bptc_via_cid.bin"), 64, 64), dec_b)


@pytest.mark.gpu
def test_facade_compress_with_stats(fh, oracle, facade_bin, tmp_path):
    """BPTCC::CompressWithStats on a sub-job and CompressImageData with a log
    stream: the blocks and the log text the oracle gives (watermark entries
    aside, which follow the process-wide counter)."""
    img = synth.generate(64, 64, 2)
    raw = tmp_path / "in.raw"
    img.tofile(raw)
    r = subprocess.run([facade_bin, "stats", str(raw), "64", "64", str(tmp_path)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    rd = lambda n: np.fromfile(tmp_path / n, dtype=np.uint8)
    for name, rect, first in (("sub", (8, 4, 40, 48), 1 * 16 + 2), ("whole", None, 0)):
        ref, _, rst = oracle.bc7_compress_logged(img, rect=rect)
        got = rd(f"stats_{name}.bin")
        assert np.array_equal(oracle.bc7_decompress(got, 64, 64),
                              oracle.bc7_decompress(ref, 64, 64)), name
        n = int(np.count_nonzero((rst["mode"] != 0) | (rst["path"] != 0)))
        want = "".join(oracle.format_block_stats(rst[k], first + k) for k in range(n))
        assert (tmp_path / f"stats_{name}.log").read_text() == want, name


@pytest.mark.gpu
def test_facade_shape_selection_function(fh, oracle, facade_bin, tmp_path):
    """CompressionSettings::m_ShapeSelectionFn called per block by the
    facade (two-subset shape (x / 4 + y / 4 + 5) % 64), against the oracle
    given the same selections (watermark entries aside)."""
    img = synth.generate(64, 64, 3)
    raw = tmp_path / "in.raw"
    img.tofile(raw)
    r = subprocess.run([facade_bin, "select", str(raw), "64", "64", str(tmp_path)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    sels = np.zeros(256, oracle.SHAPE_SELECTION_DTYPE)
    for k in range(256):
        by, bx = divmod(k, 16)
        sels[k]["num_shapes"] = 1
        sels[k]["shapes"][0] = (2, (bx + by + 5) % 64)
        sels[k]["selected_modes"] = 0xFF
    ref, _ = oracle.bc7_compress_selected(img, sels)
    got = np.fromfile(tmp_path / "select.bin", dtype=np.uint8)
    assert np.array_equal(oracle.bc7_decompress(got, 64, 64), oracle.bc7_decompress(ref, 64, 64))


@pytest.mark.gpu
def test_facade_compress_image(fh, facade_bin, tmp_path):
    """CompressImage (Core/src/TexComp.cpp:367-424) of a 62x62 image: padded
    with zeros to 64x64 (with the reference's warning), the blocks of the
    padded image, and DecompressImage's decode of them."""
    img = synth.generate(64, 64, 4)
    raw = tmp_path / "in.raw"
    img.tofile(raw)
    r = subprocess.run([facade_bin, "image", str(raw), "64", "64", str(tmp_path)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Padding with zeros" in r.stderr
    padded = np.zeros_like(img)
    padded[:62, :62] = img[:62, :62]
    want = fh.compress(padded, quality=0, wm_counter=ctypes.c_uint32(0xFFFFFFFF))
    got = np.fromfile(tmp_path / "image.bin", dtype=np.uint8)
    dec = np.fromfile(tmp_path / "image_dec.bin", dtype=np.uint32).reshape(64, 64)
    # watermark entries of solid blocks follow the process-wide counter
    assert np.array_equal(fh.decompress(got, 64, 64), fh.decompress(want, 64, 64))
    assert np.array_equal(dec, fh.decompress(got, 64, 64))


@pytest.mark.gpu
def test_facade_compress_atomic(fh, facade_bin, tmp_path):
    """BPTCC::CompressAtomic from two threads over a list of three jobs: each
    job is encoded once, with the default settings (50 annealing steps)."""
    img = synth.generate(64, 64, 5)
    raw = tmp_path / "in.raw"
    img.tofile(raw)
    r = subprocess.run([facade_bin, "atomic", str(raw), "64", "64", str(tmp_path)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    parts = [img, np.ascontiguousarray(img[:32]), np.ascontiguousarray(img[:, :32])]
    for k, part in enumerate(parts):
        h, w = part.shape
        want = fh.compress(part, quality=50, wm_counter=ctypes.c_uint32(0xFFFFFFFF))
        got = np.fromfile(tmp_path / f"atomic{k}.bin", dtype=np.uint8)
        assert np.array_equal(fh.decompress(got, w, h), fh.decompress(want, w, h)), k
