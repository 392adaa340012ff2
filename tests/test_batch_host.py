"""Host side of the batched ensemble (CPU only): the 256-instance workload
that make_instances(256, seed=0) draws, its buckets, the two initial
profiles and the per-instance configuration of the sequential comparison."""

import math

import numpy as np

from vivlasov1d_amd.batch import bucket_key, make_instances, sequential_config
from vivlasov1d_amd.core import bump_on_tail, landau


def _plain(specs):
    return [{k: v for k, v in s.items() if k != 'profile'} for s in specs]


def test_ensemble_follows_the_stated_distribution():
    specs = make_instances(256, seed=0)
    assert len(specs) == 256 and [s['index'] for s in specs] == list(range(256))
    assert {s['nx'] for s in specs} == {64, 128, 256, 512}
    assert {s['nv'] for s in specs} == {128, 256, 512, 1024}
    for s in specs:
        assert 0.2 <= s['k'] <= 0.6 and math.isclose(s['L'], 2 * math.pi / s['k'], rel_tol=1e-15)
        assert 0.005 <= s['amp'] <= 0.5
        assert s['vmax'] == 10.0 and s['ht'] == 0.05 and s['bracket'] == 'arakawaj4'
    landau_count = sum(s['profile_kind'] == 'landau' for s in specs)
    assert 128 - 40 < landau_count < 128 + 40          # probability 1/2 each (5 standard deviations)
    # the workload's size, as the benchmark reports it
    assert sum(s['nx'] * s['nv'] for s in specs) == 34193408
    # drawn from default_rng(seed): the same seed, the same ensemble
    assert _plain(make_instances(256, seed=0)) == _plain(specs)
    assert _plain(make_instances(256, seed=1)) != _plain(specs)


def test_buckets_are_the_sixteen_grid_shapes():
    specs = make_instances(256, seed=0)
    keys = {bucket_key(s) for s in specs}
    assert len(keys) == 16
    assert {(k[0], k[1]) for k in keys} == {(nx, nv) for nx in (64, 128, 256, 512) for nv in (128, 256, 512, 1024)}


def test_profiles_have_unit_mean_density():
    # n(x) = int f dv = 1 + amp cos(k x) for both profiles (the bump-on-tail
    # base 0.9 M(v) + 0.1 M((v - 4.5)/0.5)/0.5 integrates to 0.9 + 0.1)
    v = np.linspace(-12.0, 12.0, 24001)
    x = np.linspace(0.0, 10.0, 7)
    for prof in (landau(0.3, 0.2), bump_on_tail(0.3, 0.2)):
        n = np.trapezoid(prof(x[:, None], v[None, :]), v, axis=1)
        np.testing.assert_allclose(n, 1 + 0.2 * np.cos(0.3 * x), rtol=1e-12)
    bot = bump_on_tail(0.3, 0.0)
    assert bot(0.0, 4.5) > bot(0.0, 3.0)              # the bump on the tail


def test_sequential_config_matches_the_instance():
    s = make_instances(4, seed=0)[2]
    cfg = sequential_config(s, ksp_rtol=1e-9, snes_atol=1e-11, snes_max_iter=7, nt=3)
    g, sv = cfg['grid'], cfg['solver']
    assert (g['nx'], g['nv'], g['L'], g['vmax'], g['ht'], g['nt']) == (s['nx'], s['nv'], s['L'], 10.0, 0.05, 3)
    assert sv['poisson_bracket'] == 'arakawaj4' and sv['preconditioner_scheme'] == 'tensor_kinetic'
    assert (sv['petsc_ksp_rtol'], sv['petsc_snes_atol'], sv['petsc_snes_max_iter']) == (1e-9, 1e-11, 7)
