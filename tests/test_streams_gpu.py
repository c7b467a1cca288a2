"""Device entry points called on different streams share one per-device work
area and watermark counter; each call is ordered after the previous one, so
concurrent calls on two streams write what sequential calls write."""
import numpy as np
import pytest

from fastc_amd import synth
from tests.solid_blocks import nine_solid

pytestmark = pytest.mark.gpu


def test_two_streams_equal_sequential(fh):
    import torch
    imgs = [nine_solid(synth.generate(256, 256, s)) for s in (31, 32)]
    d_imgs = [torch.from_numpy(i.view(np.int32)).cuda() for i in imgs]
    size = 256 * 256
    seq = []
    for d in d_imgs:
        out = torch.zeros(size, dtype=torch.uint8, device="cuda")
        fh.compress_device(d.data_ptr(), out.data_ptr(), size, 256, 256, quality=8, seed=5)
        torch.cuda.synchronize()
        seq.append(out.cpu().numpy())
    print("two streams, two 256x256 q=8 encodes", flush=True)
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs = [torch.zeros(size, dtype=torch.uint8, device="cuda") for _ in range(2)]
    for d, o, s in zip(d_imgs, outs, streams):
        fh.compress_device(d.data_ptr(), o.data_ptr(), size, 256, 256, quality=8, seed=5,
                           stream=s.cuda_stream)
    torch.cuda.synchronize()
    for o, want in zip(outs, seq):
        assert np.array_equal(o.cpu().numpy(), want)
