"""Initial data from the run configuration (run_base.py:408-496), host only.

* the nfourier filter against scipy.fftpack, the routine the reference
  calls (rfft over x, entries >= nfourier zeroed, irfft): within 1e-14;
* a field whose x-spectrum lies inside the kept coefficients is unchanged;
* density_python / temperature_python / distribution_python modules and the
  constant density / temperature give the Maxwellian n(x) exp(-v^2/2T(x)) /
  sqrt(2 pi T(x)) of Toolbox.initialise_distribution_nT, in closed form.
"""

import math
import sys

import numpy as np
import pytest

from vivlasov1d_amd.core import Config, Grid, fourier_filter, initial_data


@pytest.mark.parametrize("nx", [100, 64, 51])
@pytest.mark.parametrize("nfourier", [1, 2, 7, 50, 200])
def test_fourier_filter_matches_fftpack(nx, nfourier):
    from scipy.fftpack import irfft, rfft
    rng = np.random.default_rng(nx + nfourier)
    f = rng.random((12, nx))                                # natural order (nv, nx)
    fT = f.T.copy()                                         # the reference's (nx, nv) view
    F = rfft(fT, axis=0)
    F[nfourier:, :] = 0.
    ref = irfft(F, axis=0).T
    got = fourier_filter(f, nfourier)
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(f).max()


def test_fourier_filter_keeps_low_modes():
    nx = 100
    x = np.arange(nx) * 2 * np.pi / nx
    f = (1 + 0.3 * np.cos(x) - 0.2 * np.sin(3 * x))[None, :] * np.ones((5, 1))
    # modes 0, 1 and 3 sit in FFTPACK entries 0..6
    assert np.abs(fourier_filter(f, 7) - f).max() <= 1e-14
    # entry 5 (Re y3) and 6 (Im y3) dropped: the sin(3x) term goes
    g = fourier_filter(f, 5)
    assert np.abs(g - (1 + 0.3 * np.cos(x))[None, :]).max() <= 1e-14
    assert fourier_filter(f, 0) is f


def _write_module(tmp_path, name, body):
    (tmp_path / (name + ".py")).write_text(body)


def test_density_and_temperature_modules(tmp_path, monkeypatch):
    _write_module(tmp_path, "vv_test_dens", "import math\n"
                  "def density(x, L):\n    return 1. + 0.01 * math.cos(0.1 * (x - 0.5 * L))\n")
    _write_module(tmp_path, "vv_test_temp", "def temperature(x):\n    return 1.0 + 0.1 * x\n")
    monkeypatch.syspath_prepend(str(tmp_path))
    cfg = Config({'grid': {'nx': 100, 'nv': 40, 'L': 20 * math.pi, 'vmin': -8.0, 'vmax': 8.0, 'ht': 0.1},
                  'initial_data': {'density_python': 'vv_test_dens', 'temperature_python': 'vv_test_temp'}})
    g = cfg.make_grid()
    f = initial_data(cfg, g)
    n = 1. + 0.01 * np.cos(0.1 * (g.x - 0.5 * g.L))
    T = 1.0 + 0.1 * g.x
    ref = n[None, :] * np.exp(-0.5 * g.v[:, None] ** 2 / T[None, :]) / np.sqrt(2 * np.pi * T[None, :])
    ref[:2] = 0
    ref[-2:] = 0
    assert np.abs(f - ref).max() <= 1e-15 * np.abs(ref).max() * 10
    # the jeans_strong.cfg filter (nfourier = 50) keeps its single mode
    cfg['initial_data']['nfourier'] = 50
    cfg['initial_data']['temperature_python'] = None
    f50 = initial_data(cfg, g)
    f0 = initial_data(Config({'grid': cfg['grid'], 'initial_data': {'density_python': 'vv_test_dens'}}), g)
    assert np.abs(f50 - f0).max() <= 1e-15
    # constants (density, temperature keys)
    cfg2 = Config({'grid': cfg['grid'], 'initial_data': {'density': 2.0, 'temperature': 0.5}})
    f2 = initial_data(cfg2, g)
    ref2 = 2.0 * np.exp(-g.v ** 2) / np.sqrt(np.pi)
    assert np.abs(f2[5] - ref2[5]).max() <= 1e-15 * 4
    sys.modules.pop("vv_test_dens", None)
    sys.modules.pop("vv_test_temp", None)


def test_distribution_module_and_missing_module(tmp_path, monkeypatch):
    _write_module(tmp_path, "vv_test_dist", "import math\n"
                  "def distribution(x, v):\n    return (1 + 0.5 * math.cos(0.5 * x)) * math.exp(-0.5 * v * v)\n")
    monkeypatch.syspath_prepend(str(tmp_path))
    cfg = Config({'grid': {'nx': 20, 'nv': 16, 'L': 4 * math.pi, 'vmin': -6.0, 'vmax': 6.0, 'ht': 0.1},
                  'initial_data': {'distribution_python': 'vv_test_dist'}})
    g = cfg.make_grid()
    f = initial_data(cfg, g)
    ref = (1 + 0.5 * np.cos(0.5 * g.x))[None, :] * np.exp(-0.5 * g.v[:, None] ** 2)
    ref[:2] = 0
    ref[-2:] = 0
    assert np.abs(f - ref).max() <= 1e-15
    sys.modules.pop("vv_test_dist", None)
    cfg['initial_data']['distribution_python'] = 'vv_no_such_module'
    with pytest.raises(ImportError, match="vv_no_such_module"):
        initial_data(cfg, g)


def test_unknown_preconditioner_scheme_runs_without_one():
    """TensorProductPreconditioner.create returns None for names other than
    the four tensor schemes (examples/twostream_bgk*.cfg use 'TensorFast')."""
    from vivlasov1d_amd.solvers import effective_preconditioner
    for s in ('tensor_kinetic', 'tensor_kinetic_scipy', 'tensor_potential', 'tensor_potential_scipy', ''):
        assert effective_preconditioner(s) == s
    cfg = Config({'solver': {'preconditioner_type': 'tensor', 'preconditioner_scheme': 'TensorFast'}})
    with pytest.warns(UserWarning, match="without a preconditioner"):
        assert effective_preconditioner(cfg.get_preconditioner()) == ''
    with pytest.warns(UserWarning):
        assert effective_preconditioner('tensor_diagonal') == ''
