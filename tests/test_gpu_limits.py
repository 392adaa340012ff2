"""The largest accepted grid lengths and the context's memory ledger.

* nv = 4096 (kinetic) and nx = 4096 (potential) are the longest tridiagonal
  systems the library accepts: the solve then runs 256 chunks of 16 rows,
  whose chunk buffer (256 x 17 complex) plus the reduced-system arrays take
  about 84 KiB of LDS per workgroup.  P^-1 (FP64, FP32 and transposed) at
  16 x 4096 and 4096 x 16 against the oracle restatement: <= 1e-12 relative
  in FP64, <= 2e-6 in FP32, dot-product identity of P^-T to 1e-12.
* The FFT's longest rows, nx = 4096 (kinetic), are covered by the same case
  run the other way round (4096 x 16).
* vv_scratch_bytes grows by what a first-use buffer adds (the FP32 basis of
  the mixed solve, the explicit-guess stages), and is never below the Krylov
  basis alone.
"""

import numpy as np
import pytest

from oracle import vlasov_oracle as O

from test_gpu_parity import _nat, _setup

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nx,nv,precond", [(16, 4096, 'tensor_kinetic'), (4096, 16, 'tensor_kinetic'),
                                           (4096, 16, 'tensor_potential'), (16, 4096, 'tensor_potential')])
def test_longest_systems_match_oracle(nx, nv, precond, cuda):
    import torch
    print("case", nx, nv, precond, flush=True)
    g, ctx, H, rng, T = _setup(nx, nv, 'arakawaj4', 11 + nx, cuda, precond=precond)
    if precond == 'tensor_potential':
        x = np.arange(nx) * 2 * np.pi / nx
        phi = 0.3 * np.cos(x) + 0.05 * rng.standard_normal(nx)
        ctx.precond_update(T(phi))
        pc = O.PotentialPreconditioner(g, phi, 'scipy')
    else:
        pc = O.Preconditioner(g, 'scipy')
    x = rng.standard_normal((nx, nv))
    ref = pc.apply(x)
    y = ctx.new_field()
    ctx.precond(T(_nat(x)), y)
    err = np.abs(y.cpu().numpy().T - ref).max() / np.abs(ref).max()
    assert err <= 1e-12, err
    x32 = x.astype(np.float32).astype(np.float64)
    ref32 = pc.apply(x32)
    xd = T(_nat(x32)).float()
    y32 = torch.empty_like(xd)
    ctx.precond32(xd, y32)
    err32 = np.abs(y32.double().cpu().numpy().T - ref32).max() / np.abs(ref32).max()
    assert err32 <= 2e-6, err32
    u, v = T(rng.standard_normal((nv, nx))), T(rng.standard_normal((nv, nx)))
    Pu, PTv = ctx.new_field(), ctx.new_field()
    ctx.precond(u, Pu)
    ctx.precond_t(v, PTv)
    lhs, rhs = (v * Pu).sum().item(), (PTv * u).sum().item()
    assert abs(lhs - rhs) <= 1e-12 * (v.abs() * Pu.abs()).sum().item(), (lhs, rhs)


def test_scratch_bytes_count_first_use_buffers(cuda):
    import math
    from vivlasov1d_amd.core import Config
    from vivlasov1d_amd.runner import Runscript
    prof = lambda x, v: (1 + 0.1 * np.cos(0.5 * x)) * np.exp(-0.5 * v ** 2) / math.sqrt(2 * math.pi)
    cfg = Config({'grid': {'nx': 32, 'nv': 64, 'L': 4 * math.pi, 'vmax': 8.0, 'ht': 0.05, 'nt': 1},
                  'solver': {'poisson_bracket': 'arakawaj4', 'preconditioner_type': 'tensor',
                             'preconditioner_scheme': 'tensor_kinetic', 'initial_guess': 'symplectic2'},
                  'initial_data': {'charge': -1.0}})
    run = Runscript(cfg, profile=prof)
    N = 32 * 64
    b0 = run.ctx.scratch_bytes()
    assert b0 >= 8 * 31 * N
    run.step()                                  # explicit guess: 6 N doubles on first use
    b1 = run.ctx.scratch_bytes()
    assert b1 - b0 >= 6 * 8 * N, (b0, b1)
    mixed = Runscript(cfg, profile=prof, precision='mixed')
    m0 = mixed.ctx.scratch_bytes()
    mixed.step()                                # + FP32 basis (31 vectors) and the FP64 residual
    m1 = mixed.ctx.scratch_bytes()
    assert m1 - m0 >= 4 * 31 * N + 8 * N, (m0, m1)
