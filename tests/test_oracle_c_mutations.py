"""The C restatement's pins catch plausible mistakes (CPU only).

Each case compiles a copy of oracle/c/vv1d_oracle.c with one mistake planted
(a bracket sign, the J4 weight, the J2 normalisation, a GMRES rotation sign,
the real-FFT split, a preconditioner band, the H_ave composition, the
residual's average) into a temporary directory, loads it in place of the
built library, and requires that at least two of the comparisons with the
NumPy oracle in test_oracle_c.py fail."""

import subprocess

import numpy as np
import pytest

import test_oracle_c as TC
from oracle import build as B
from oracle import cbaseline as CB

# (text in the C source, its replacement)
CASES = {
    'J4 weight sign': ('(2. * (s1 * r->af1) - s2 * r->af2)', '(2. * (s1 * r->af1) + s2 * r->af2)'),
    'J1 jcp sign': ('+ x_pm * (h_pc - h_cm);', '- x_pm * (h_pc - h_cm);'),
    'J2 normalisation': ('r->af2 = r->hx_inv * r->hv_inv / 24.;', 'r->af2 = r->hx_inv * r->hv_inv / 12.;'),
    'GMRES rotation sign': ('h[j] = cs[j - 1] * h[j] - sn[j - 1] * t0;', 'h[j] = cs[j - 1] * h[j] + sn[j - 1] * t0;'),
    'real-FFT split sign': ('-0.5 * I * (zk - zm)', '0.5 * I * (zk - zm)'),
    'sub-diagonal from the upper difference': ('((er * 0.5) * gm * af) + I * ((ei * 0.5) * gm * af)',
                                               '((er * 0.5) * gp * af) + I * ((ei * 0.5) * gp * af)'),
    'H_ave composition index': ('0.5 * r->hp3[i]) + 0.5 * r->hp4[i]', '0.5 * r->hp2[i]) + 0.5 * r->hp4[i]'),
    'residual average weight': ('0.5 * fh[(long)(jj) * nx + (ii)] + 0.5 * f[', 'fh[(long)(jj) * nx + (ii)] + 0.5 * f['),
}


def _pins():
    return [lambda k=k: TC.test_c_operators_match_oracle(CB, k, 16, 40) for k in ('arakawaj1', 'arakawaj2', 'arakawaj4')] + \
           [lambda n=n: TC.test_c_precond_at_workload_lengths(CB, n) for n in (64, 256)] + \
           [lambda s=s: TC.test_c_timesteps_match_oracle(CB, *s) for s in ((32, 64), (25, 48))]


def _failures(pins):
    failed = 0
    for c in pins:
        try:
            with np.errstate(all='ignore'):
                c()
        except AssertionError:
            failed += 1
    return failed


def _build(src, path):
    # the recipe's flags at -O1 (contraction stays off)
    subprocess.run(["gcc", "-O1", "-std=gnu11", "-march=x86-64-v3", "-ffp-contract=off", "-fopenmp", "-fPIC",
                    "-shared", "-o", str(path), str(src), "-lm"], check=True)


@pytest.mark.parametrize("name", sorted(CASES))
def test_planted_c_mistake_fails_a_pin(name, tmp_path, monkeypatch):
    old, new = CASES[name]
    text = open(B.SRC).read()
    assert text.count(old) == 1, name
    src = tmp_path / "planted.c"
    src.write_text(text.replace(old, new))
    lib = tmp_path / "libplanted.so"
    _build(src, lib)
    monkeypatch.setattr(CB, 'LIB', str(lib))
    monkeypatch.setattr(CB, '_lib', None)
    assert _failures(_pins()) >= 2, name


def test_unplanted_c_source_passes_every_pin(tmp_path, monkeypatch):
    lib = tmp_path / "libunplanted.so"
    _build(B.SRC, lib)
    monkeypatch.setattr(CB, 'LIB', str(lib))
    monkeypatch.setattr(CB, '_lib', None)
    assert _failures(_pins()) == 0
