"""HDF5 output in the reference's schema (run_base.py:528-562, 766-783), on
the CPU: round trip through the writer and the reader, the layout read back by
the HDF5 library's own h5dump (an independent reader), and the accesses the
reference's readers make (Grid.pyx:53-59, distribution_function.pyx:319-324,
hamiltonian.pyx:159-166, potential.pyx:64-66)."""

import os
import shutil
import subprocess

import numpy as np
import pytest

from vivlasov1d_amd import hdf5 as H

try:
    H.library()
    HAVE = True
except H.HDF5Error:
    HAVE = False

pytestmark = pytest.mark.skipif(not HAVE, reason="no HDF5 C library in this image")


def _write(path, nx=8, nv=6, steps=3, seed=0):
    rng = np.random.default_rng(seed)
    x = np.arange(nx) * 0.5
    v = -1.0 + np.arange(nv) * 0.4
    data = []
    with H.HDF5Output(path, x, v, -1.0) as w:
        for k in range(steps):
            d = dict(t=0.1 * k, h0=rng.random((nv, nx)), h1=rng.random((nv, nx)), h2=np.zeros((nv, nx)),
                     f=rng.random((nv, nx)), phi_int=rng.random(nx), phi_ext=np.zeros(nx), n=rng.random(nx),
                     u=rng.random(nx), e=rng.random(nx))
            w.save(**d)
            data.append(d)
    return x, v, data


def test_round_trip(tmp_path):
    p = str(tmp_path / "run.hdf5")
    x, v, data = _write(p)
    d = H.read(p)
    assert d["attrs"]["charge"] == -1.0
    assert np.array_equal(d["x"], x) and np.array_equal(d["v"], v)
    assert d["t"].shape == (3, 1, 1)
    for name in H.TIME_SERIES_2D:
        assert d[name].shape == (3, 6, 8)
    for name in H.TIME_SERIES_1D:
        assert d[name].shape == (3, 8)
    for k, dk in enumerate(data):
        assert d["t"][k, 0, 0] == dk["t"]
        for name in H.TIME_SERIES_2D + H.TIME_SERIES_1D:
            assert np.array_equal(d[name][k], dk[name]), (k, name)


def test_reference_reader_accesses(tmp_path):
    """What Grid.load_from_hdf5 and the diagnostics read from the file."""
    p = str(tmp_path / "run.hdf5")
    x, v, data = _write(p, nx=16, nv=10, steps=4)
    d = H.read(p)
    t = d["t"][:, 0, 0]
    assert len(t) - 1 == 3 and abs((t[1] - t[0]) - 0.1) < 1e-15
    assert abs((d["x"][1] - d["x"][0]) - 0.5) < 1e-15 and abs((d["v"][1] - d["v"][0]) - 0.4) < 1e-15
    f2 = d["f"][2, :, :].T               # (nx, nv): f[i, j] as the reference indexes it
    assert f2.shape == (16, 10)
    assert f2[3, 7] == data[2]["f"][7, 3]
    assert d["phi_int"][1, :].shape == (16,)


def test_h5dump_reads_the_layout(tmp_path):
    exe = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)
    if exe is None:
        pytest.skip("no h5dump")
    p = str(tmp_path / "run.hdf5")
    _write(p)
    head = subprocess.run([exe, "-H", p], capture_output=True, text=True, check=True).stdout
    assert 'ATTRIBUTE "charge"' in head
    assert "( 3, 6, 8 ) / ( H5S_UNLIMITED, 6, 8 )" in head          # f, h0, h1, h2
    assert "( 3, 8 ) / ( H5S_UNLIMITED, 8 )" in head                # phi_int, ..., e
    assert "( 3, 1, 1 ) / ( H5S_UNLIMITED, 1, 1 )" in head          # t
    assert "SIMPLE { ( 8 ) / ( 8 ) }" in head and "SIMPLE { ( 6 ) / ( 6 ) }" in head
    tval = subprocess.run([exe, "-d", "/t", "-y", "-w", "0", p], capture_output=True, text=True, check=True).stdout
    body = tval[tval.index("DATA {") + 6:tval.index("}", tval.index("DATA {"))]
    assert [float(s) for s in body.replace(",", " ").split()] == [0.0, 0.1, 0.2]
