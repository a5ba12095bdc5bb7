"""msx_multislice_noalloc (the slice loop for callers that cannot allocate
device memory), called through the C ABI: its exit waves equal
msx_multislice's on the same device buffers, on two calls in a row (the
second on a smaller grid with fewer waves and layers, so it runs on what the
first left in the shared scratch array), a repeated call changes no device
memory accounting, and a stack whose tables exceed the scratch array is
refused before anything runs."""
import ctypes as C

import numpy as np
import pytest
import torch

import multislice_amd as msx
from multislice_amd import engine

pytestmark = pytest.mark.gpu


def _case(seed, n, nl, ns, nwave, dtype):
    rng = np.random.default_rng(seed)
    L = 0.1 * n
    V = torch.from_numpy((np.abs(rng.standard_normal((nl, n, n))).cumsum(axis=1) % 25.0).astype(
        np.float64 if dtype == torch.complex128 else np.float32)).cuda()
    waves = torch.from_numpy((rng.standard_normal((nwave, n, n)) + 1j * rng.standard_normal((nwave, n, n)))
                             .astype(np.complex128 if dtype == torch.complex128 else np.complex64)).cuda()
    return dict(V=V, n=n, L=L, cz=rng.uniform(1.5, 2.5, nl), layer=rng.integers(0, nl, ns), waves=waves)


def _noalloc(ctx, trans, cz, layer, psi, L, kev, prec, psi0=None):
    st, keep = engine.Context._stack(psi.shape[-2], psi.shape[-1], L, L, kev, cz, layer)
    torch.cuda.synchronize()
    rc = msx.load().msx_multislice_noalloc(ctx.h, C.byref(st), C.c_void_p(trans.data_ptr()),
                                           C.c_void_p(psi0.data_ptr()) if psi0 is not None else None,
                                           C.c_void_p(psi.data_ptr()), psi.shape[0], prec)
    return rc


@pytest.mark.parametrize("dtype,prec", [(torch.complex64, msx.F32), (torch.complex128, msx.F64)])
def test_noalloc_equals_multislice_twice(dtype, prec):
    ctx = msx.Context(0)
    for seed, n, nl, ns, nwave, kev in ((1, 512, 6, 23, 5, 300.0), (2, 256, 3, 9, 2, 200.0)):
        c = _case(seed, n, nl, ns, nwave, dtype)
        t = ctx.transmission(c["V"], c["L"], c["L"], kev)
        ref = c["waves"].clone()
        ctx.multislice(t, c["cz"], c["layer"], ref, c["L"], c["L"], kev)
        got = c["waves"].clone()
        assert _noalloc(ctx, t, c["cz"], c["layer"], got, c["L"], kev, prec) == 0, msx.load().msx_last_error()
        assert torch.equal(got, ref), (n, prec)
        # with a shared incident wave
        ref0 = torch.empty_like(ref)
        ctx.multislice(t, c["cz"], c["layer"], ref0, c["L"], c["L"], kev, psi0=c["waves"][0])
        got0 = torch.empty_like(ref)
        assert _noalloc(ctx, t, c["cz"], c["layer"], got0, c["L"], kev, prec, psi0=c["waves"][0]) == 0
        assert torch.equal(got0, ref0)
    ctx.close()


def test_noalloc_allocates_nothing_and_refuses_large_stacks():
    ctx = msx.Context(0)
    c = _case(3, 512, 4, 12, 3, torch.complex64)
    t = ctx.transmission(c["V"], c["L"], c["L"], 300.0)
    psi = c["waves"].clone()
    assert _noalloc(ctx, t, c["cz"], c["layer"], psi, c["L"], 300.0, msx.F32) == 0  # loads the kernels
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(3):
        psi = c["waves"].clone()
        torch.cuda.synchronize()
        f0, _ = torch.cuda.mem_get_info()
        assert _noalloc(ctx, t, c["cz"], c["layer"], psi, c["L"], 300.0, msx.F32) == 0
        f1, _ = torch.cuda.mem_get_info()
        assert f1 == f0
    # 511 layers at 512^2 FP32 need more than MSX_NOALLOC_SCRATCH_BYTES (510 fit)
    big = np.full(511, 2.0)
    rc = _noalloc(ctx, t, big, np.zeros(4, np.int64), psi, c["L"], 300.0, msx.F32)
    assert rc == -4 and b"scratch" in msx.load().msx_last_error()
    assert _noalloc(ctx, t, np.full(510, 2.0), np.zeros(4, np.int64), psi, c["L"], 300.0, msx.F32) == 0
    ctx.close()
