"""Pins of the oracle's Arakawa brackets, residual and Jacobian (CPU only).

Each check is independent of the oracle's own formula:
  * the reference's second statement of the brackets (the assembled
    stencil tables of jacobian(), transcribed in oracle/stencil_tables.py);
  * Arakawa's identities (antisymmetry, conservation of sum f, sum f^2 and
    sum h weights);
  * closed-form discrete symbols for f = cos(kx) g(v), h = h(v), derived by
    hand from the stencil definition;
  * consistency order: J1 and J2 are 2nd-order, J4 = 2 J1 - J2 is 4th-order
    accurate for {f, h} = f_x h_v - f_v h_x (a wrong J4 weight drops it to 2nd);
  * the residual is affine in f, so (F(f+e) - F(f-e))/2 = J e exactly.
"""

import numpy as np
import pytest

from oracle import stencil_tables as ST
from oracle import vlasov_oracle as O

KINDS = ['arakawaj1', 'arakawaj2', 'arakawaj4']


def _grid(nx=8, nv=16, L=2 * np.pi, vmax=3.0, ht=0.1):
    return O.Grid(nx, nv, L, vmax, ht=ht)


@pytest.mark.parametrize("kind", KINDS)
def test_matrix_free_equals_stencil_table(kind):
    g = _grid(8, 16)
    rng = np.random.default_rng(0)
    H = rng.standard_normal((g.nx, g.nv))
    M = ST.bracket_matrix(kind, H, g, factor=0.7)
    N = g.nx * g.nv
    for col in range(N):
        e = np.zeros(N)
        e[col] = 1.0
        x = e.reshape((g.nv, g.nx)).T
        y = O.bracket(kind, x, H, g, factor=0.7)
        np.testing.assert_allclose(y.T.reshape(N), M[:, col], rtol=0, atol=1e-12 * np.abs(M).max())


@pytest.mark.parametrize("kind", KINDS)
def test_antisymmetry(kind):
    g = _grid(16, 32)
    rng = np.random.default_rng(1)
    f, h = rng.standard_normal((2, g.nx, g.nv))
    a = O.bracket(kind, f, h, g)
    b = O.bracket(kind, h, f, g)
    assert np.abs(a + b).max() <= 1e-13 * np.abs(a).max()


@pytest.mark.parametrize("kind", KINDS)
def test_conservation(kind):
    """sum B = sum f B = sum h B = 0 for fields vanishing within 4 rows of the
    v-boundaries (no stencil term then touches the unbracketed rows)."""
    g = _grid(16, 40)
    rng = np.random.default_rng(2)
    f, h = rng.standard_normal((2, g.nx, g.nv))
    f[:, :4] = f[:, -4:] = 0.0
    h[:, :4] = h[:, -4:] = 0.0
    B = O.bracket(kind, f, h, g)
    s = np.abs(B).sum()
    assert abs(B.sum()) <= 1e-13 * s
    assert abs((f * B).sum()) <= 1e-13 * np.abs(f * B).sum()
    assert abs((h * B).sum()) <= 1e-13 * np.abs(h * B).sum()


def _symbol_case(kind):
    g = _grid(32, 24, L=2 * np.pi, vmax=2.0)
    k = 3.0
    rng = np.random.default_rng(3)
    gv = rng.standard_normal(g.nv)
    hv = rng.standard_normal(g.nv)
    f = np.cos(k * g.x)[:, None] * gv[None, :]
    h = np.broadcast_to(hv[None, :], (g.nx, g.nv)).copy()
    return g, k, gv, hv, f, h


@pytest.mark.parametrize("kind", KINDS)
def test_closed_form_symbol(kind):
    """h = h(v), f = cos(k x) g(v): hand-derived stencil sums
    J1: S1(x) [2 g_j (h+ - h-) + g_{j+1}(h+ - h0) + g_{j-1}(h0 - h-)] / (12 hx hv)
    J2: [S1(x)(g_{j+1} + g_{j-1})(h+ - h-) + S2(x) g_j (h+ - h-)
         + S1(x)(g_{j+1}(h++ - h0) + g_{j-1}(h0 - h--))] / (24 hx hv)
    with S1 = -2 sin(k hx) sin(k x), S2 = -2 sin(2 k hx) sin(k x)."""
    g, k, gv, hv, f, h = _symbol_case(kind)
    B = O.bracket(kind, f, h, g)
    S1 = -2 * np.sin(k * g.hx) * np.sin(k * g.x)[:, None]
    S2 = -2 * np.sin(2 * k * g.hx) * np.sin(k * g.x)[:, None]
    j = np.arange(2, g.nv - 2)
    G = lambda d: gv[j + d][None, :]
    Hh = lambda d: hv[j + d][None, :]
    j1 = S1 * (2 * G(0) * (Hh(1) - Hh(-1)) + G(1) * (Hh(1) - Hh(0)) + G(-1) * (Hh(0) - Hh(-1))) / (12 * g.hx * g.hv)
    j2 = (S1 * (G(1) + G(-1)) * (Hh(1) - Hh(-1)) + S2 * G(0) * (Hh(1) - Hh(-1))
          + S1 * (G(1) * (Hh(2) - Hh(0)) + G(-1) * (Hh(0) - Hh(-2)))) / (24 * g.hx * g.hv)
    ref = {'arakawaj1': j1, 'arakawaj2': j2, 'arakawaj4': 2 * j1 - j2}[kind]
    got = B[:, 2:g.nv - 2]
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.all(B[:, :2] == 0) and np.all(B[:, -2:] == 0)


def _order(kind, ns):
    errs = []
    for n in ns:
        g = O.Grid(n, n + 1, 2 * np.pi, np.pi, vmin=-np.pi)
        X, V = np.meshgrid(g.x, g.v, indexing='ij')
        f = np.sin(X) * np.cos(V) + 0.3 * np.cos(2 * X + V)
        h = np.cos(X) * np.sin(2 * V) + 0.5 * V ** 2
        fx = np.cos(X) * np.cos(V) - 0.6 * np.sin(2 * X + V)
        fv = -np.sin(X) * np.sin(V) - 0.3 * np.sin(2 * X + V)
        hx = -np.sin(X) * np.sin(2 * V)
        hv = 2 * np.cos(X) * np.cos(2 * V) + V
        exact = fx * hv - fv * hx
        B = O.bracket(kind, f, h, g)
        errs.append(np.abs(B - exact)[:, 2:-2].max())
    return np.log2(np.array(errs[:-1]) / np.array(errs[1:]))


@pytest.mark.parametrize("kind,order", [('arakawaj1', 2), ('arakawaj2', 2), ('arakawaj4', 4)])
def test_consistency_order(kind, order):
    rates = _order(kind, [48, 96, 192])
    assert rates[-1] > order - 0.25 and rates[-1] < order + 0.5, rates


@pytest.mark.parametrize("kind", KINDS)
def test_residual_affine_and_jacobian(kind):
    g = _grid(8, 16, ht=0.05)
    rng = np.random.default_rng(4)
    f = rng.random((g.nx, g.nv))
    fh = rng.random((g.nx, g.nv))
    H = rng.standard_normal((g.nx, g.nv))
    for col in rng.choice(g.nx * g.nv, 12, replace=False):
        e = np.zeros(g.nx * g.nv)
        e[col] = 1.0
        e = e.reshape((g.nv, g.nx)).T
        fd = (O.residual(kind, f + e, fh, H, g) - O.residual(kind, f - e, fh, H, g)) / 2.0
        Je = O.jacobian(kind, e, H, g)
        assert np.abs(fd - Je).max() <= 1e-12 * np.abs(Je).max()


@pytest.mark.parametrize("kind", KINDS)
def test_residual_structure(kind):
    g = _grid(8, 16, ht=0.05)
    rng = np.random.default_rng(5)
    f, fh = rng.random((2, g.nx, g.nv))
    H = rng.standard_normal((g.nx, g.nv))
    F = O.residual(kind, f, fh, H, g)
    # boundary rows carry only the time derivative
    np.testing.assert_array_equal(F[:, :2], (f - fh)[:, :2] * g.ht_inv)
    np.testing.assert_array_equal(F[:, -2:], (f - fh)[:, -2:] * g.ht_inv)
    # f = fh: pure bracket of f with factor 1
    F0 = O.residual(kind, f, f, H, g)
    np.testing.assert_allclose(F0, O.bracket(kind, f, H, g), rtol=0, atol=1e-13 * np.abs(F0).max())
    # Jacobian of a constant-in-time state: u/dt + 1/2 B(u, H)
    J = O.jacobian(kind, f, H, g)
    np.testing.assert_allclose(J, f / g.ht + 0.5 * O.bracket(kind, f, H, g), rtol=1e-14, atol=1e-12)


def test_h_average_order_and_values():
    rng = np.random.default_rng(6)
    h0 = rng.standard_normal(16)
    h1p, h1h, h2p, h2h = rng.standard_normal((4, 8))
    H = O.h_average(h0, h1p, h1h, h2p, h2h)
    exp = h0[None, :] + 0.5 * (h1p + h1h + h2p + h2h)[:, None]
    np.testing.assert_allclose(H, exp, rtol=1e-15, atol=1e-15)
    # element (i, j) is exactly ((((h0_j + .5 h1p_i) + .5 h1h_i) + .5 h2p_i) + .5 h2h_i)
    i, j = 3, 5
    assert H[i, j] == ((((h0[j] + .5 * h1p[i]) + .5 * h1h[i]) + .5 * h2p[i]) + .5 * h2h[i])
