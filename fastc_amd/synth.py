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
Every random choice is keyed by (seed, block index) or (seed, pixel), so a
sub-rectangle of a larger image is stable.

Pixels are packed as uint32 with R in the low byte (the reference's
`Pixel::Pack`, Base/src/Pixel.cpp), i.e. bytes R, G, B, A in memory.
"""
import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        z = x
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def _hash(*parts):
    h = _splitmix64(np.uint64(0x5EED5EED5EED5EED))
    for p in parts:
        h = _splitmix64(h ^ np.asarray(p, dtype=np.uint64))
    return h


# 16 ramp directions: round(64*cos, 64*sin) for angles k*22.5 degrees.
_DIRS = np.array([(64, 0), (59, 24), (45, 45), (24, 59), (0, 64), (-24, 59),
                  (-45, 45), (-59, 24), (-64, 0), (-59, -24), (-45, -45),
                  (-24, -59), (0, -64), (24, -59), (45, -45), (59, -24)],
                 dtype=np.int64)


def _value_noise(seed, ch, width, height, x0=0, y0=0):
    ys = np.arange(y0, y0 + height, dtype=np.int64)[:, None]
    xs = np.arange(x0, x0 + width, dtype=np.int64)[None, :]
    acc = np.zeros((height, width), dtype=np.int64)
    for octave in range(4):
        shift = 5 - octave                       # spacing 32, 16, 8, 4
        amp = 8 >> octave                        # persistence 0.5
        lx = xs >> shift
        ly = ys >> shift
        tx = (xs & ((1 << shift) - 1)) << (16 - shift)   # Q16 in [0, 1)
        ty = (ys & ((1 << shift) - 1)) << (16 - shift)
        sx = (tx * tx * (3 * 65536 - 2 * tx)) >> 32        # smoothstep, Q16
        sy = (ty * ty * (3 * 65536 - 2 * ty)) >> 32

        def lat(ax, ay):
            return (_hash(seed, 1 + ch, octave, ax.astype(np.uint64) & _M64,
                          ay.astype(np.uint64) & _M64) & np.uint64(0xFF)).astype(np.int64)

        l00 = lat(lx, ly)
        l10 = lat(lx + 1, ly)
        l01 = lat(lx, ly + 1)
        l11 = lat(lx + 1, ly + 1)
        top = (l00 << 16) + (l10 - l00) * sx               # Q16
        bot = (l01 << 16) + (l11 - l01) * sx
        v = (top << 16) + (bot - top) * sy                 # Q32
        acc += amp * v
    # acc / (15 * 2^32) is in [0, 255]
    return acc // (15 << 32)


def generate(width, height, seed, edges=False, x0=0, y0=0):
    """Return a (height, width) uint32 array of packed RGBA8 pixels.

    width/height must be multiples of 4.  x0/y0 place the tile inside a larger
    virtual image (the result is identical to cropping that image)."""
    if width % 4 or height % 4:
        raise ValueError("width and height must be multiples of 4")
    ys = np.arange(y0, y0 + height, dtype=np.int64)[:, None]
    xs = np.arange(x0, x0 + width, dtype=np.int64)[None, :]
    pix_id = ((ys.astype(np.uint64) << np.uint64(32)) | xs.astype(np.uint64))
    rgb = []
    for ch in range(3):
        v = _value_noise(seed, ch, width, height, x0, y0)
        noise = (_hash(seed, 10 + ch, pix_id) % np.uint64(17)).astype(np.int64) - 8
        rgb.append(np.clip(v + noise, 0, 255))
    r, g, b = rgb

    bx = xs >> 2
    by = ys >> 2
    blk_id = ((by.astype(np.uint64) << np.uint64(32)) | bx.astype(np.uint64))
    blk_id = np.broadcast_to(blk_id, (height, width))
    xl = (xs & 3)
    yl = (ys & 3)

    # Alpha: 75% opaque blocks, 25% directional ramps 0..255 across the block.
    aroll = (_hash(seed, 20, blk_id) % np.uint64(100)).astype(np.int64)
    adir = (_hash(seed, 21, blk_id) % np.uint64(16)).astype(np.int64)
    dx = _DIRS[adir, 0]
    dy = _DIRS[adir, 1]
    proj = dx * (2 * xl - 3) + dy * (2 * yl - 3)
    # min/max of the projection over the 4x4 block: corners are extremal
    corners = np.stack([dx * cx + dy * cy for cx in (-3, 3) for cy in (-3, 3)])
    pmin = corners.min(axis=0)
    pmax = corners.max(axis=0)
    ramp = ((proj - pmin) * 255) // np.maximum(pmax - pmin, 1)
    a = np.where(aroll < 75, 255, ramp)

    # 10% of blocks become one flat colour.
    froll = (_hash(seed, 30, blk_id) % np.uint64(100)).astype(np.int64)
    fcol = _hash(seed, 31, blk_id)
    fr = (fcol & np.uint64(0xFF)).astype(np.int64)
    fg = ((fcol >> np.uint64(8)) & np.uint64(0xFF)).astype(np.int64)
    fb = ((fcol >> np.uint64(16)) & np.uint64(0xFF)).astype(np.int64)
    fa = np.where(((fcol >> np.uint64(24)) % np.uint64(4)) == 0,
                  ((fcol >> np.uint64(32)) & np.uint64(0xFF)).astype(np.int64), 255)
    flat = froll < 10
    r = np.where(flat, fr, r)
    g = np.where(flat, fg, g)
    b = np.where(flat, fb, b)
    a = np.where(flat, fa, a)

    if edges:
        # 5% of the non-flat blocks: two colours split by a line through the block.
        eroll = (_hash(seed, 40, blk_id) % np.uint64(100)).astype(np.int64)
        e = (eroll < 5) & ~flat
        c0 = _hash(seed, 41, blk_id)
        c1 = _hash(seed, 42, blk_id)
        edir = (_hash(seed, 43, blk_id) % np.uint64(16)).astype(np.int64)
        eoff = (_hash(seed, 44, blk_id) % np.uint64(5)).astype(np.int64) - 2
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
