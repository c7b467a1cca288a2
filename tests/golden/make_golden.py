"""Regenerate the golden fixtures in tests/golden/ from the reference.

Run in a container that has the FasTC tree at /root/reference and the built
oracle/_ref library (make -C oracle/ref):

    python tests/golden/make_golden.py [--tables-only]

Fixtures written (data only, no reference source text):
  bc7_tables.json  SHA-256 of the reference's BC7 tables as flat arrays: the
                   two lookup tables (BPTCEncoder/src/BCLookupTables.h:
                   Optimal7CompressBC7Mode5, Optimal6CompressDXT1, uint8), the
                   partition masks (Shapes.h, uint16), anchors
                   (AnchorTables.h, uint8) and interpolation weights
                   (Compressor.cpp kInterpolationValues, uint32).
  dxt_etc.npz      DXT1 / DXT5 / ETC1 output of the reference encoders
                   (stb_dxt.h, rg_etc1.cpp built from the reference tree) for a
                   64x64 generator image (seed 2, hard edges).
  dxt_etc_hashes.json  SHA-256 of the same encoders' output for the 256x256
                   and 2048x2048 seed-2 images (the DXT workload's full size).
  decode.npz       the reference's rg_etc1::unpack_etc1_block (built from the
                   reference tree) on 256 random 8-byte blocks (64x64,
                   numpy seed 11) and on the 64x64 ETC1 fixture above.

    python tests/golden/make_golden.py --decode-only   rewrites decode.npz only
"""
import ctypes
import hashlib
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = os.environ.get("FASTC_REF", "/root/reference")


def parse_table(text, name, dtype=np.uint8):
    i = text.index(name)
    body = text[text.index("=", i):text.index("};", i)]
    return np.array([int(x, 0) for x in re.findall(r"0x[0-9a-fA-F]+|\d+", body)], dtype=dtype)


def partition_tables(ref):
    """The reference's partition masks, anchors and interpolation weights as
    flat arrays (Shapes.h, AnchorTables.h, Compressor.cpp)."""
    shapes = open(os.path.join(ref, "BPTCEncoder", "include", "FasTC", "Shapes.h")).read()
    anchors = open(os.path.join(ref, "BPTCEncoder", "src", "AnchorTables.h")).read()
    comp = open(os.path.join(ref, "BPTCEncoder", "src", "Compressor.cpp")).read()
    return {
        "shape2": parse_table(shapes, "kShapeMask2[kNumShapes2]", np.uint16),
        "shape3": parse_table(shapes, "kShapeMask3[kNumShapes3][2]", np.uint16),
        "anchor2": parse_table(anchors, "kAnchorIdx2[BPTCC::kNumShapes2]"),
        "anchor3": parse_table(anchors, "kAnchorIdx3[2][BPTCC::kNumShapes3]"),
        "interp": parse_table(comp, "kInterpolationValues[4][16][2]", np.uint32),
    }


def ref_lib():
    path = os.path.join(ROOT, "oracle", "_ref", "libfastc_ref.so")
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    L.ref_dxt_compress.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, vp]
    L.ref_etc1_compress.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, vp]
    L.ref_etc1_decompress.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, vp]
    return L


def encode_all(L, img):
    h, w = img.shape
    n = (h // 4) * (w // 4)
    d1 = np.zeros(n * 8, np.uint8)
    d5 = np.zeros(n * 16, np.uint8)
    e1 = np.zeros(n * 8, np.uint8)
    L.ref_dxt_compress(img.ctypes.data, w, h, 0, d1.ctypes.data)
    L.ref_dxt_compress(img.ctypes.data, w, h, 1, d5.ctypes.data)
    L.ref_etc1_compress(img.ctypes.data, w, h, e1.ctypes.data)
    return d1, d5, e1


def write_decode_fixture(L):
    from fastc_amd import synth
    rnd = np.random.default_rng(11).integers(0, 256, 64 * 64 // 2, dtype=np.uint8)
    enc = encode_all(L, synth.generate(64, 64, 2, edges=True))[2]
    out = {"etc1_random": rnd, "etc1_encoded": enc}
    for key, name in (("etc1_random", "etc1_random_decoded"), ("etc1_encoded", "etc1_decoded")):
        dec = np.zeros((64, 64), np.uint32)
        L.ref_etc1_decompress(out[key].ctypes.data, 64, 64, dec.ctypes.data)
        out[name] = dec
    np.savez_compressed(os.path.join(HERE, "decode.npz"), **out)


def main():
    from fastc_amd import synth
    text = open(os.path.join(REF, "BPTCEncoder", "src", "BCLookupTables.h")).read()
    m5 = parse_table(text, "Optimal7CompressBC7Mode5[256][2]")
    d6 = parse_table(text, "Optimal6CompressDXT1[256][2][3]")
    assert m5.size == 512 and d6.size == 1536
    tables = {"mode5_sha256": hashlib.sha256(m5.tobytes()).hexdigest(),
              "dxt6_sha256": hashlib.sha256(d6.tobytes()).hexdigest()}
    for name, arr in partition_tables(REF).items():
        tables[name + "_sha256"] = hashlib.sha256(arr.tobytes()).hexdigest()
    with open(os.path.join(HERE, "bc7_tables.json"), "w") as f:
        json.dump(tables, f, indent=1)
    if "--tables-only" in sys.argv:
        print("table fixture written")
        return

    L = ref_lib()
    write_decode_fixture(L)
    if "--decode-only" in sys.argv:
        print("decode fixture written")
        return
    img = synth.generate(64, 64, 2, edges=True)
    d1, d5, e1 = encode_all(L, img)
    np.savez_compressed(os.path.join(HERE, "dxt_etc.npz"), image=img, dxt1=d1, dxt5=d5, etc1=e1)
    hashes = {}
    for size in (256, 2048):
        img = synth.generate(size, size, 2, edges=True)
        d1, d5, e1 = encode_all(L, img)
        hashes[str(size)] = {"dxt1": hashlib.sha256(d1.tobytes()).hexdigest(),
                             "dxt5": hashlib.sha256(d5.tobytes()).hexdigest(),
                             "etc1": hashlib.sha256(e1.tobytes()).hexdigest()}
    with open(os.path.join(HERE, "dxt_etc_hashes.json"), "w") as f:
        json.dump(hashes, f, indent=1)
    print("fixtures written")


if __name__ == "__main__":
    main()
