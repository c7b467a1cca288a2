"""Mip levels and padding on the GPU (fh_mip_level_device, fh_pad4_device)
against their definition, spelled out in numpy: per channel the rounded mean
of the 2x2 (2x1, 1x2) source texels, sides halving rounding down; padding to
whole 4x4 blocks with transparent black (Core/src/TexComp.cpp:382-393)."""
import numpy as np
import pytest
import torch

from fastc_amd import batch_workload, synth

pytestmark = pytest.mark.gpu


def next_level(img):
    h, w = img.shape
    c = img.view(np.uint8).reshape(h, w, 4).astype(np.int64)
    dh, dw = max(1, h // 2), max(1, w // 2)
    out = np.zeros((dh, dw, 4), np.int64)
    for y in range(dh):
        for x in range(dw):
            if w > 1 and h > 1:
                s = c[2 * y, 2 * x] + c[2 * y, 2 * x + 1] + c[2 * y + 1, 2 * x] + c[2 * y + 1, 2 * x + 1]
                out[y, x] = (s + 2) // 4
            elif w > 1:
                out[y, x] = (c[0, 2 * x] + c[0, 2 * x + 1] + 1) // 2
            else:
                out[y, x] = (c[2 * y, 0] + c[2 * y + 1, 0] + 1) // 2
    return np.ascontiguousarray(out.astype(np.uint8)).view(np.uint32).reshape(dh, dw)


def gpu_level(fh, img):
    h, w = img.shape
    src = torch.from_numpy(np.ascontiguousarray(img).view(np.int32)).cuda()
    dh, dw = max(1, h // 2), max(1, w // 2)
    dst = torch.zeros(dh * dw, dtype=torch.int32, device="cuda")
    fh.mip_level_device(src.data_ptr(), w, h, dst.data_ptr())
    torch.cuda.synchronize()
    return dst.cpu().numpy().view(np.uint32).reshape(dh, dw)


@pytest.mark.parametrize("w,h", [(64, 32), (5, 3), (1, 8), (8, 1), (2, 2), (33, 17)])
def test_mip_level(fh, w, h):
    rng = np.random.default_rng(w * 100 + h)
    img = rng.integers(0, 2 ** 32, (h, w), dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(gpu_level(fh, img), next_level(img))


def test_mip_chain_matches_workload_generator(fh):
    """The chain the texture-batch workload builds on the CPU, level by level."""
    img = synth.generate(64, 16, 12)
    want = batch_workload.mip_chain(img)
    cur = img
    for level in want[1:]:
        cur = gpu_level(fh, cur)
        assert np.array_equal(cur, level)
    assert cur.shape == (1, 1)


@pytest.mark.parametrize("w,h", [(1, 1), (2, 3), (5, 9), (8, 8), (13, 4)])
def test_pad4(fh, w, h):
    img = np.random.default_rng(w + 7 * h).integers(1, 2 ** 32, (h, w), dtype=np.uint64).astype(np.uint32)
    src = torch.from_numpy(img.view(np.int32)).cuda()
    ph, pw = (h + 3) // 4 * 4, (w + 3) // 4 * 4
    dst = torch.full((ph * pw,), -1, dtype=torch.int32, device="cuda")
    fh.pad4_device(src.data_ptr(), w, h, dst.data_ptr())
    torch.cuda.synchronize()
    got = dst.cpu().numpy().view(np.uint32).reshape(ph, pw)
    want = np.zeros((ph, pw), np.uint32)
    want[:h, :w] = img
    assert np.array_equal(got, want)
    assert np.array_equal(batch_workload.pad4(img), want)


def test_one_by_one_has_no_next_level(fh):
    src = torch.zeros(1, dtype=torch.int32, device="cuda")
    with pytest.raises(fh.FastHipError):
        fh.mip_level_device(src.data_ptr(), 1, 1, src.data_ptr())
