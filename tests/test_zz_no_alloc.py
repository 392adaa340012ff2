"""vv_moments_noalloc (the allocation-free variant of vv_moments, scratch in a
file-scope device array) against vv_moments through the C ABI: n, phi, h1
and the Poisson residual norm bit for bit, on two calls in a row with the
second on a smaller grid (the shared array keeps the first call's contents),
for a single-instance and a batched context; the context's device-memory
ledger does not move; nx * batch > 4096 is rejected with VV_ERR_INVALID."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ctx(nx, nv, hx=None):
    from vivlasov1d_amd.core import Grid
    from vivlasov1d_amd.solvers import Context
    g = Grid(nx, nv, 4 * np.pi, 8.0, ht=0.05)
    return Context(g, bracket='arakawaj4', charge=-1.0, hx=hx)


def _check(ctx, rng):
    import torch
    shape = (ctx.B, ctx.grid.nv, ctx.grid.nx) if ctx.batched else (ctx.grid.nv, ctx.grid.nx)
    f = torch.from_numpy(rng.random(shape)).cuda()
    ref = [ctx.new_x() for _ in range(3)]
    pn_ref = np.atleast_1d(ctx.moments(f, *ref, want_norm=True))
    got = [ctx.new_x() for _ in range(3)]
    pn = torch.full((ctx.B,), -1.0, dtype=torch.float64, device='cuda')
    before = ctx.scratch_bytes()
    ctx.moments_noalloc(f, *got, pnorm=pn)
    torch.cuda.synchronize()
    assert ctx.scratch_bytes() == before
    for a, b in zip(got, ref):
        assert torch.equal(a, b)
    assert np.array_equal(pn.cpu().numpy(), pn_ref), (pn.cpu().numpy(), pn_ref)
    # without a norm output
    again = [ctx.new_x() for _ in range(3)]
    ctx.moments_noalloc(f, *again)
    for a, b in zip(again, ref):
        assert torch.equal(a, b)


def test_moments_noalloc_matches_moments(cuda):
    rng = np.random.default_rng(5)
    print("case noalloc 256x512 then 64x128", flush=True)
    _check(_ctx(256, 512), rng)
    _check(_ctx(64, 128), rng)          # smaller: the array still holds the larger call's data
    print("case noalloc batched 3 x (128x64) then 2 x (32x16)", flush=True)
    _check(_ctx(128, 64, hx=[4 * np.pi / 128, 2 * np.pi / 0.3 / 128, 10 * np.pi / 128]), rng)
    _check(_ctx(32, 16, hx=[4 * np.pi / 32, 10 * np.pi / 32]), rng)


def test_moments_noalloc_rejects_oversize(cuda):
    import torch
    from vivlasov1d_amd import _lib
    ctx = _ctx(4096, 8, hx=[4 * np.pi / 4096] * 2)
    f = ctx.new_field()
    xs = [ctx.new_x() for _ in range(3)]
    with pytest.raises(_lib.VVError, match="4096"):
        ctx.moments_noalloc(f, *xs)
    torch.cuda.synchronize()
