"""msx_copy_probe (measurement only): the contiguous copy and the copies with
the column sweep's access pattern move every value exactly once, and
malformed requests are refused before any launch."""
import numpy as np
import pytest

import multislice_amd as msx

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode,nimg,ny", [(0, 3, 512), (0, 1, 8), (8, 3, 512), (16, 2, 256), (8, 1, 64), (1008, 3, 512),
                                         (1008, 1, 64)])
def test_copy_probe_copies(mode, nimg, ny):
    import torch
    ctx = msx.Context(0)
    g = torch.Generator(device="cuda").manual_seed(mode + ny)
    src = torch.randn((nimg, 512, ny), dtype=torch.complex64, device="cuda", generator=g)
    dst = torch.zeros_like(src)
    ms = ctx.copy_probe(mode, src, dst, nimg, 512, ny)
    assert ms > 0
    assert bool((dst == src).all())
    ctx.close()


def test_copy_probe_rejects():
    import torch
    ctx = msx.Context(0)
    a = torch.zeros((1, 512, 512), dtype=torch.complex64, device="cuda")
    b = torch.zeros_like(a)
    for mode, nx, ny in ((8, 256, 512), (16, 512, 504), (4, 512, 512), (0, 512, 4), (1008, 512, 500)):
        with pytest.raises(msx.MsxError):
            ctx.copy_probe(mode, a, b, 1, nx, ny)
    ctx.close()
