"""Test helper: images whose solid-block count is a multiple of 9, so that
encodes sharing the process-wide watermark counter (the reference's gWMVal,
9 table entries) start at the same entry."""
import numpy as np


def nine_solid(img):
    """Break solid 4x4 blocks (one pixel nudged) until their count is a
    multiple of 9."""
    img = img.copy()
    h, w = img.shape
    solid = [(y, x) for y in range(0, h, 4) for x in range(0, w, 4)
             if (img[y:y + 4, x:x + 4] == img[y, x]).all()]
    for y, x in solid[:len(solid) % 9]:
        img[y, x] ^= np.uint32(1)
    return img
