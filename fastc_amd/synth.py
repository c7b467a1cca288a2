"""Seeded synthetic RGBA8 test images for the block-encode workloads.

The reference ships no benchmark images, so workloads are generated.  All
arithmetic is integer (numpy uint64/int64), so the same seed yields the same
bytes on every machine.

Image model (per the benchmark definition):
  * R, G, B: independent 4-octave value noise.  Lattice spacing 32 px halving
    per octave, persistence 0.5 (amplitudes 8:4:2:1), smoothstep (3t^2-2t^3)
    interpolation in Q16 fixed point, mapped to 0..255, plus uniform integer
    noise in [-8, 8], clamped.
  * A: 255 in 75% of 4x4 blocks; in the rest a linear 0..255 ramp across the
    block along one of 16 directions.
  * 10% of blocks are replaced by one flat RGBA colour (the solid path).
  * Optionally (the DXT workload) 5% of blocks become two-colour hard edges.
Every random choice is keyed by (seed, lattice point), (seed, block) or
(seed, pixel) in absolute image coordinates, so a tile of a larger image is
the same as the crop of that image.

Pixels are packed as uint32 with R in the low byte (the reference's
`Pixel::Pack`, Base/src/Pixel.cpp), i.e. bytes R, G, B, A in memory.
"""
import os
import threading

import numpy as np


def _splitmix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def _hash(*parts):
    h = _splitmix64(np.uint64(0x5EED5EED5EED5EED))
    for p in parts:
        h = _splitmix64(h ^ np.asarray(p, dtype=np.uint64))
    return h


def _xy_key(ys, xs):
    """(y, x) integer grids -> uint64 keys (y << 32 | x)."""
    return (ys.astype(np.uint64) << np.uint64(32)) | xs.astype(np.uint64)


# 16 ramp directions: round(64*cos, 64*sin) for angles k*22.5 degrees.
_DIRS = np.array([(64, 0), (59, 24), (45, 45), (24, 59), (0, 64), (-24, 59),
                  (-45, 45), (-59, 24), (-64, 0), (-59, -24), (-45, -45),
                  (-24, -59), (0, -64), (24, -59), (45, -45), (59, -24)],
                 dtype=np.int64)


def _value_noise(seed, ch, width, height, x0, y0):
    ys = np.arange(y0, y0 + height, dtype=np.int64)[:, None]
    xs = np.arange(x0, x0 + width, dtype=np.int64)[None, :]
    acc = np.zeros((height, width), dtype=np.int64)
    for octave in range(4):
        shift = 5 - octave                       # spacing 32, 16, 8, 4
        amp = 8 >> octave                        # persistence 0.5
        lx0 = x0 >> shift
        ly0 = y0 >> shift
        nlx = ((x0 + width - 1) >> shift) - lx0 + 2
        nly = ((y0 + height - 1) >> shift) - ly0 + 2
        gy = np.arange(ly0, ly0 + nly, dtype=np.int64)[:, None]
        gx = np.arange(lx0, lx0 + nlx, dtype=np.int64)[None, :]
        lat = (_hash(seed, 1 + ch, octave, _xy_key(gy, gx)) & np.uint64(0xFF)).astype(np.int64)
        lx = (xs >> shift) - lx0
        ly = (ys >> shift) - ly0
        tx = (xs & ((1 << shift) - 1)) << (16 - shift)   # Q16 in [0, 1)
        ty = (ys & ((1 << shift) - 1)) << (16 - shift)
        sx = (tx * tx * (3 * 65536 - 2 * tx)) >> 32        # smoothstep, Q16
        sy = (ty * ty * (3 * 65536 - 2 * ty)) >> 32
        l00 = lat[ly, lx]
        l10 = lat[ly, lx + 1]
        l01 = lat[ly + 1, lx]
        l11 = lat[ly + 1, lx + 1]
        top = (l00 << 16) + (l10 - l00) * sx               # Q16
        bot = (l01 << 16) + (l11 - l01) * sx
        acc += amp * ((top << 16) + (bot - top) * sy)      # Q32
    # acc / (15 * 2^32) is in [0, 255]
    return acc // (15 << 32)


_SYNTH_LIB = None


def _synth_lib():
    global _SYNTH_LIB
    if _SYNTH_LIB is None:
        import ctypes
        import os
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsynth.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path)
        L.synth_generate.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.c_uint64, ctypes.c_int, ctypes.c_int64,
                                     ctypes.c_int64]
        L.synth_generate.restype = ctypes.c_int
        _SYNTH_LIB = L
    return _SYNTH_LIB


def generate(width, height, seed, edges=False, x0=0, y0=0, impl="auto"):
    """Return a (height, width) uint32 array of packed RGBA8 pixels.

    width/height (and x0/y0) must be multiples of 4.  x0/y0 place the tile
    inside a larger virtual image (the result equals cropping that image).
    impl: "numpy" (this module's definition), "c" (libsynth.so, the same
    arithmetic in C) or "auto" (C when built)."""
    if impl in ("auto", "c"):
        L = _synth_lib()
        if L is None and impl == "c":
            raise ImportError("libsynth.so is not built")
        if L is not None:
            out = np.empty((height, width), dtype=np.uint32)
            if width % 4 or height % 4 or x0 % 4 or y0 % 4:
                raise ValueError("sizes and offsets must be multiples of 4")
            # row bands in parallel (the C call releases the GIL); each band
            # is the crop of the same virtual image, so the result does not
            # depend on the split
            nthreads = max(1, min(16, os.cpu_count() or 1, height // 64))
            rows = height // 4
            bands = [(4 * (rows * i // nthreads), 4 * (rows * (i + 1) // nthreads))
                     for i in range(nthreads)]
            status = [0] * nthreads

            def work(i):
                r0, r1 = bands[i]
                if r1 > r0:
                    status[i] = L.synth_generate(out[r0:r1].ctypes.data, width, r1 - r0, seed,
                                                 int(edges), x0, y0 + r0)

            if nthreads == 1:
                work(0)
            else:
                th = [threading.Thread(target=work, args=(i,)) for i in range(nthreads)]
                for t in th:
                    t.start()
                for t in th:
                    t.join()
            if any(status):
                raise ValueError("bad generator arguments")
            return out
    return generate_numpy(width, height, seed, edges, x0, y0)


def generate_numpy(width, height, seed, edges=False, x0=0, y0=0):
    """The reference definition of generate() in numpy."""
    if width % 4 or height % 4 or x0 % 4 or y0 % 4:
        raise ValueError("sizes and offsets must be multiples of 4")
    ys = np.arange(y0, y0 + height, dtype=np.int64)[:, None]
    xs = np.arange(x0, x0 + width, dtype=np.int64)[None, :]
    noise = _hash(seed, 10, _xy_key(ys, xs))
    rgb = []
    for ch in range(3):
        v = _value_noise(seed, ch, width, height, x0, y0)
        nz = (((noise >> np.uint64(16 * ch)) & np.uint64(0xFFFF)) % np.uint64(17)).astype(np.int64) - 8
        rgb.append(np.clip(v + nz, 0, 255))
    r, g, b = rgb

    # per-block random choices, computed at block resolution then expanded
    bh, bw = height // 4, width // 4
    by = np.arange(y0 // 4, y0 // 4 + bh, dtype=np.int64)[:, None]
    bx = np.arange(x0 // 4, x0 // 4 + bw, dtype=np.int64)[None, :]
    bkey = _xy_key(by, bx)
    hb1 = _hash(seed, 20, bkey)
    hb2 = _hash(seed, 30, bkey)

    def up(a):
        return np.repeat(np.repeat(a, 4, axis=0), 4, axis=1)

    xl = (xs & 3)
    yl = (ys & 3)

    # Alpha: 75% opaque blocks, 25% directional ramps 0..255 across the block.
    aroll = up((hb1 % np.uint64(100)).astype(np.int64))
    adir = up(((hb1 >> np.uint64(8)) % np.uint64(16)).astype(np.int64))
    dx = _DIRS[adir, 0]
    dy = _DIRS[adir, 1]
    proj = dx * (2 * xl - 3) + dy * (2 * yl - 3)
    # extremes of the projection over a 4x4 block are at its corners
    corners = np.stack([dx * cx + dy * cy for cx in (-3, 3) for cy in (-3, 3)])
    pmin = corners.min(axis=0)
    pmax = corners.max(axis=0)
    ramp = ((proj - pmin) * 255) // np.maximum(pmax - pmin, 1)
    a = np.where(aroll < 75, 255, ramp)

    # 10% of blocks become one flat colour (alpha 255 in 3 of 4 of them).
    froll = (hb2 % np.uint64(100)).astype(np.int64)
    fcol = _hash(seed, 31, bkey)
    flat = up(froll < 10)
    fr = up((fcol & np.uint64(0xFF)).astype(np.int64))
    fg = up(((fcol >> np.uint64(8)) & np.uint64(0xFF)).astype(np.int64))
    fb = up(((fcol >> np.uint64(16)) & np.uint64(0xFF)).astype(np.int64))
    fa = up(np.where(((fcol >> np.uint64(24)) % np.uint64(4)) == 0,
                     ((fcol >> np.uint64(32)) & np.uint64(0xFF)).astype(np.int64), 255))
    r = np.where(flat, fr, r)
    g = np.where(flat, fg, g)
    b = np.where(flat, fb, b)
    a = np.where(flat, fa, a)

    if edges:
        # 5% of the non-flat blocks: two colours split by a line through the block.
        he = _hash(seed, 40, bkey)
        e = up((he % np.uint64(100)).astype(np.int64) < 5) & ~flat
        c0 = up(_hash(seed, 41, bkey))
        c1 = up(_hash(seed, 42, bkey))
        edir = up(((he >> np.uint64(8)) % np.uint64(16)).astype(np.int64))
        eoff = up(((he >> np.uint64(16)) % np.uint64(5)).astype(np.int64)) - 2
        side = (_DIRS[edir, 0] * (2 * xl - 3) + _DIRS[edir, 1] * (2 * yl - 3)) > eoff * 32
        sel = np.where(side, c1, c0)
        r = np.where(e, (sel & np.uint64(0xFF)).astype(np.int64), r)
        g = np.where(e, ((sel >> np.uint64(8)) & np.uint64(0xFF)).astype(np.int64), g)
        b = np.where(e, ((sel >> np.uint64(16)) & np.uint64(0xFF)).astype(np.int64), b)

    out = (r.astype(np.uint32) | (g.astype(np.uint32) << 8) |
           (b.astype(np.uint32) << 16) | (a.astype(np.uint32) << 24))
    return np.ascontiguousarray(out, dtype=np.uint32)


def as_bytes(img):
    """Packed uint32 image -> (h, w, 4) uint8 view, bytes R, G, B, A."""
    return img.view(np.uint8).reshape(img.shape[0], img.shape[1], 4)


def blocks(img):
    """(h, w) packed image -> (h/4 * w/4, 16) blocks in raster order, the
    gather of the reference's GetBlock (BPTCEncoder/src/Compressor.cpp:1460)."""
    h, w = img.shape
    return (img.reshape(h // 4, 4, w // 4, 4).transpose(0, 2, 1, 3)
            .reshape(-1, 16).copy())


def unblocks(blk, width, height):
    """Inverse of blocks()."""
    return (blk.reshape(height // 4, width // 4, 4, 4).transpose(0, 2, 1, 3)
            .reshape(height, width).copy())
