"""The Arnoldi step's fused kernels: irFFT + Gram-Schmidt dots
(irfft_mdot_kernel) and the Jacobian-vector product + rFFT (jac_rfft_kernel).

Where the FFT workgroup's rows are one CGS slice with the CGS kernels' own
vectors per thread (256x512 and 1024x2048 in FP64; 1024x2048 in FP32 with
cgs_ept = 4, where the FP32 J u's tile offsets of H follow the bracket
kernel's 64 x 16 tiles), the fused launches must reproduce jacobian + rfft
and irfft + mdot exactly: same iterations, same converged reason, equal solutions (the
fused rFFT's first radix-4 stage multiplies by the unit twiddle, which can
change only the sign of an exact zero), for single and batched contexts, for
J1, J2 and J4 and 2, 4 and 8 rows per FFT workgroup.  At 512x1024 the fused
slice holds 2 vectors per thread where mdot's holds 4, so only the summation
tree of the dots differs: same iteration count, solutions within 1e-12
relative.  The profile counters show which path ran.  (The GMRES parity
tests against the oracle run the fused path at 256x512, the default.)
"""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ctx(nx, nv, cuda, hx=None, restart=30, bracket='arakawaj4'):
    import torch
    from vivlasov1d_amd.core import Grid, kinetic_hamiltonian
    from vivlasov1d_amd.solvers import Context
    g = Grid(nx, nv, 4 * math.pi, 8.0, ht=0.05)
    s = torch.cuda.current_stream().cuda_stream
    ctx = Context(g, bracket=bracket, charge=-1.0, restart=restart, hx=hx, stream=s)
    B = ctx.B
    rng = np.random.default_rng(nx + nv)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    h1 = [0.05 * rng.standard_normal((B, nx) if B > 1 else nx) for _ in range(4)]
    ctx.set_hamiltonian(T(kinetic_hamiltonian(g)), *(T(h) for h in h1))
    shape = (B, nv, nx) if B > 1 else (nv, nx)
    f = rng.random(shape)
    f[..., :2, :] = 0
    f[..., -2:, :] = 0
    fh = f + 1e-3 * rng.standard_normal(shape)
    b = torch.zeros(shape, dtype=torch.float64, device=cuda)
    ctx.function(T(f), T(fh), b)            # b = P^-1 F(f)
    return ctx, b


def _solve(ctx, b, fused, mixed=False, counts=None):
    import torch
    ctx.set_fusion(fused)
    ctx.profile(("irfft_mdot", "irfft_mdot32", "mdot", "mdot32", "jac_rfft", "rfft", "jac_rfft32", "rfft32"))
    x = torch.zeros_like(b)
    # a random right-hand side: 120 iterations (four restart cycles) whether
    # or not they reach rtol, the comparison is of the two paths
    if mixed:
        res, _ = ctx.ksp_solve_mixed(b, x, 1e-10, 1e-50, 1e5, 120)
    else:
        res = ctx.ksp_solve(b, x, 1e-10, 1e-50, 1e5, 120)
    ctx.synchronize()
    n_fused = ctx.profile_read("irfft_mdot32" if mixed else "irfft_mdot")[1]
    n_mdot = ctx.profile_read("mdot32" if mixed else "mdot")[1]
    if counts is not None:
        sfx = "32" if mixed else ""
        counts["jac_rfft"] = ctx.profile_read("jac_rfft" + sfx)[1]
        counts["rfft"] = ctx.profile_read("rfft" + sfx)[1]
    ctx.profile(())
    return res, x.cpu().numpy(), n_fused, n_mdot


def _its(res):
    return [(r.its, r.reason) for r in res] if isinstance(res, list) else (res.its, res.reason)


@pytest.mark.parametrize("nx,nv", [(256, 512), (1024, 2048)])
def test_fused_fp64_bitwise(nx, nv, cuda):
    print("case fused fp64", nx, nv, flush=True)
    ctx, b = _ctx(nx, nv, cuda)
    c1, c0 = {}, {}
    r1, x1, nf1, nm1 = _solve(ctx, b, True, counts=c1)
    r0, x0, nf0, nm0 = _solve(ctx, b, False, counts=c0)
    assert nf1 > 0 and nm1 == 0, (nf1, nm1)
    assert nf0 == 0 and nm0 > 0, (nf0, nm0)
    # every Arnoldi step's J v + rFFT fused (the rFFTs left are the restart residuals')
    assert c1["jac_rfft"] == nf1 and c0["jac_rfft"] == 0, (c1, c0)
    assert c1["rfft"] < c0["rfft"], (c1, c0)
    assert r1.reason in (2, -3) and _its(r1) == _its(r0)
    assert r1.rnorm == r0.rnorm
    assert np.array_equal(x1, x0)


@pytest.mark.parametrize("bracket,rows", [("arakawaj1", 2), ("arakawaj2", 4), ("arakawaj4", 8)])
def test_fused_jac_rfft_kinds_and_rows(bracket, rows, cuda):
    print("case fused jac_rfft 256x512", bracket, rows, flush=True)
    ctx, b = _ctx(256, 512, cuda, bracket=bracket)
    # CGS slices of the fused irFFT workgroup's size (rows x 256 doubles), so
    # that mdot and the fused dots sum in the same order
    ept = rows * 256 // (256 * 2)
    ctx.set_tuning(rows, ept)
    assert ctx.get_tuning() == (rows, ept)
    c1 = {}
    r1, x1, nf1, _ = _solve(ctx, b, True, counts=c1)
    r0, x0, nf0, _ = _solve(ctx, b, False)
    assert nf1 > 0 and nf0 == 0 and c1["jac_rfft"] == nf1, (nf1, nf0, c1)
    assert r1.reason in (2, -3) and _its(r1) == _its(r0)
    assert r1.rnorm == r0.rnorm
    assert np.array_equal(x1, x0)


def test_fused_jac_rfft_ragged_rows(cuda):
    # odd nv: the last FFT workgroup's second row is past the grid, and the
    # fused irFFT + dots does not apply (rows per workgroup must divide nv),
    # so J v + rFFT is fused alone, before the separate irFFT and dots
    print("case fused jac_rfft 256x509", flush=True)
    ctx, b = _ctx(256, 509, cuda)
    c1, c0 = {}, {}
    r1, x1, nf1, nm1 = _solve(ctx, b, True, counts=c1)
    r0, x0, nf0, nm0 = _solve(ctx, b, False, counts=c0)
    assert nf1 == 0 and nf0 == 0 and nm1 > 0 and nm1 == nm0, (nf1, nf0, nm1, nm0)
    assert c1["jac_rfft"] == nm1 and c0["jac_rfft"] == 0, (c1, c0)
    assert r1.reason in (2, -3) and _its(r1) == _its(r0)
    assert r1.rnorm == r0.rnorm
    assert np.array_equal(x1, x0)


def test_fused_fp64_other_slice(cuda):
    print("case fused fp64 512x1024", flush=True)
    ctx, b = _ctx(512, 1024, cuda)
    r1, x1, nf1, _ = _solve(ctx, b, True)
    r0, x0, nf0, _ = _solve(ctx, b, False)
    assert nf1 > 0 and nf0 == 0
    assert r1.reason in (2, -3) and _its(r1) == _its(r0)
    assert np.abs(x1 - x0).max() <= 1e-12 * np.abs(x0).max()


def test_fused_fp32_bitwise(cuda):
    print("case fused fp32 1024x2048", flush=True)
    ctx, b = _ctx(1024, 2048, cuda)
    ctx.set_tuning(0, 4)         # CGS slices of 4 float4 per thread: the fused slice
    c1, c0 = {}, {}
    r1, x1, nf1, nm1 = _solve(ctx, b, True, mixed=True, counts=c1)
    r0, x0, nf0, nm0 = _solve(ctx, b, False, mixed=True, counts=c0)
    assert nf1 > 0 and nm1 == 0 and nf0 == 0 and nm0 > 0
    assert c1["jac_rfft"] == nf1 and c0["jac_rfft"] == 0 and c0["rfft"] > c1["rfft"], (c1, c0)
    assert r1.reason in (2, -3) and _its(r1) == _its(r0)
    assert np.array_equal(x1, x0)


def test_fused_batched_bitwise(cuda):
    print("case fused batched 3 x 256x512", flush=True)
    Ls = [4 * math.pi, 2 * math.pi / 0.3, 10 * math.pi]
    ctx, b = _ctx(256, 512, cuda, hx=[L / 256 for L in Ls])
    c1 = {}
    r1, x1, nf1, _ = _solve(ctx, b, True, counts=c1)
    r0, x0, nf0, _ = _solve(ctx, b, False)
    assert nf1 > 0 and nf0 == 0 and c1["jac_rfft"] == nf1
    assert _its(r1) == _its(r0) and all(r.reason in (2, -3) for r in r1)
    assert np.array_equal(x1, x0)
