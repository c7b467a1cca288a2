"""Synthetic texture batch: textures of uneven power-of-two sizes with full
box-filtered mip chains and a mix of target formats.

    textures = make_batch(count=512, seed=4)

Each texture's side lengths are drawn independently as 2^k, k uniform in
[log2(min_side), log2(max_side)] (log-uniform, non-square allowed); a fraction
`rgb_only` of the textures have alpha 255 everywhere; formats are drawn with
probabilities `mix` (BPTC, DXT5, ETC1).  Level 0 comes from the seeded
generator (fastc_amd.synth, seed derived from (seed, texture index)); every
further level is the rounded 2x2 box filter of the previous one (2x1 or 1x2
once a side reaches 1), down to 1x1.  Levels narrower or shorter than 4
pixels are padded with transparent black to a multiple of 4, as
CompressImage pads (Core/src/TexComp.cpp:382-393).
"""
import numpy as np

from fastc_amd import synth

FMT_DXT5, FMT_ETC1, FMT_BPTC = 1, 2, 3


def box_filter(img):
    """One mip step of a (h, w) packed RGBA8 image."""
    h, w = img.shape
    a = img.view(np.uint8).reshape(h, w, 4).astype(np.uint32)
    if h > 1 and w > 1:
        s = a[0::2, 0::2] + a[1::2, 0::2] + a[0::2, 1::2] + a[1::2, 1::2]
        out = (s + 2) >> 2
    elif w > 1:
        out = (a[:, 0::2] + a[:, 1::2] + 1) >> 1
    else:
        out = (a[0::2, :] + a[1::2, :] + 1) >> 1
    return np.ascontiguousarray(out.astype(np.uint8)).view(np.uint32).reshape(out.shape[:2])


def pad4(img):
    """Pads a level to multiples of 4 with transparent black."""
    h, w = img.shape
    ph, pw = (h + 3) // 4 * 4, (w + 3) // 4 * 4
    if (ph, pw) == (h, w):
        return img
    out = np.zeros((ph, pw), dtype=np.uint32)
    out[:h, :w] = img
    return out


def mip_chain(level0):
    levels = [level0]
    while levels[-1].shape != (1, 1):
        levels.append(box_filter(levels[-1]))
    return levels


def make_batch(count=512, seed=4, min_side=64, max_side=2048, rgb_only=0.4,
               mix=(0.6, 0.25, 0.15)):
    """List of textures: dicts with format, size, rgb_only and `levels`
    (padded (h, w) uint32 arrays, level 0 first)."""
    rng = np.random.default_rng(seed)
    k0, k1 = int(np.log2(min_side)), int(np.log2(max_side))
    textures = []
    for i in range(count):
        w = 1 << int(rng.integers(k0, k1 + 1))
        h = 1 << int(rng.integers(k0, k1 + 1))
        opaque = bool(rng.random() < rgb_only)
        fmt = [FMT_BPTC, FMT_DXT5, FMT_ETC1][int(rng.choice(3, p=list(mix)))]
        img = synth.generate(w, h, (seed << 20) + i)
        if opaque:
            img = img | np.uint32(0xFF000000)
        levels = [pad4(lv) for lv in mip_chain(img)]
        textures.append({"format": fmt, "width": w, "height": h, "rgb_only": opaque,
                         "levels": levels})
    return textures


def texel_count(textures):
    """Texels of all levels (unpadded sizes)."""
    n = 0
    for t in textures:
        w, h = t["width"], t["height"]
        while True:
            n += w * h
            if w == 1 and h == 1:
                break
            w, h = max(1, w // 2), max(1, h // 2)
    return n
