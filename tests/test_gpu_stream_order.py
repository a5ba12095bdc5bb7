"""Device inputs produced on torch's current stream are consumed in order by
the engine, whose work runs on its own non-blocking stream (msx_wait_stream
through Context._order).  Each test queues a long chain of matrix products on
torch's stream, then the kernel that writes the engine's input, then the
engine call: without the ordering the engine would read the input before it
was written."""
import numpy as np
import pytest
import torch

import multislice_amd as msx

pytestmark = pytest.mark.gpu


def _busy():
    """About 20-60 ms of queued work on torch's current stream."""
    a = torch.full((4096, 4096), 1e-3, device="cuda")
    for _ in range(20):
        a = a @ a
    return a


def test_fft2d_waits_for_the_producer():
    ctx = msx.Context(0)
    g = torch.Generator(device="cuda").manual_seed(3)
    src = torch.randn((16, 512, 512), dtype=torch.complex64, device="cuda", generator=g)
    ref = src.clone()
    ctx.fft2d(ref, +1)
    torch.cuda.synchronize()
    x = torch.zeros_like(src)
    torch.cuda.synchronize()
    _busy()
    x.copy_(src)            # queued behind the products on torch's stream
    ctx.fft2d(x, +1)        # must see src, not zeros
    torch.cuda.synchronize()
    assert torch.equal(x, ref)
    ctx.close()


def test_displace_out_none_waits_for_its_clone():
    ctx = msx.Context(0)
    at = np.zeros(200_000, msx.ATOM_DTYPE)
    at["Z"] = 8
    at["occ"] = 1.0
    at["x"] = np.linspace(0, 100, len(at))
    host = msx.displace(at, {8: 0.1}, 11, 3)
    dev = torch.from_numpy(at.view(np.uint8)).cuda()
    torch.cuda.synchronize()
    _busy()
    out = msx.displace(dev, {8: 0.1}, 11, 3, ctx=ctx)  # out=None: cloned on torch's stream, then written
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(msx.ATOM_DTYPE)
    for k in ("x", "y", "z"):  # device and host agree to FP64 rounding; undisplaced atoms would be ~0.1 A off
        assert np.abs(got[k] - host[k]).max() < 1e-12, k
    ctx.close()
