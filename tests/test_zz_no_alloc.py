"""fh_compress_device_noalloc: the BPTC device encode in the library's fixed
per-device work area, with no allocation.  Its output must equal
fh_compress_device's bit for bit, on two calls in a row (the second smaller,
so stale contents of the shared area would show).

Both entry points advance the process-wide watermark counter (the
reference's gWMVal); the images hold a multiple of 9 solid blocks so that
each call starts at the same table entry as the call it is compared with.
"""
import ctypes

import numpy as np
import pytest

from fastc_amd import synth
from tests.solid_blocks import nine_solid

pytestmark = pytest.mark.gpu


def encode(fh, entry, img, quality, seed):
    import torch
    h, w = img.shape
    size = fh.compressed_size(fh.eCompressionFormat_BPTC, w, h)
    d_img = torch.from_numpy(img.view(np.int32)).cuda()
    out = torch.full((size,), 0x5A, dtype=torch.uint8, device="cuda")
    j = fh._job(fh.eCompressionFormat_BPTC, d_img.data_ptr(), out.data_ptr(), size, w, h, None,
                quality, seed, 0xFF, None)
    status = entry(ctypes.byref(j), ctypes.c_void_p(0))
    assert status == 0, fh.lib().fh_last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("quality", [0, 8])
def test_noalloc_equals_device_entry(fh, quality):
    L = fh.lib()
    for (w, h, seed) in ((512, 512, 21), (128, 64, 22)):
        print(f"noalloc vs device: {w}x{h} q={quality} seed={seed}", flush=True)
        img = nine_solid(synth.generate(w, h, seed))
        want = encode(fh, L.fh_compress_device, img, quality, 77)
        got = encode(fh, L.fh_compress_device_noalloc, img, quality, 77)
        bad = np.nonzero((got.reshape(-1, 16) != want.reshape(-1, 16)).any(axis=1))[0]
        assert bad.size == 0, f"{bad.size} blocks differ, first {bad[:8]}"
