"""ctypes access to the C restatement (libvv1d_oracle.so): the host-core CPU
baseline of bench.py and a second checker for tests.  Checker only."""

import ctypes
import os
import time

import numpy as np

from . import vlasov_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libvv1d_oracle.so")
BRACKETS = {'arakawaj1': 1, 'arakawaj2': 2, 'arakawaj4': 4}
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle C library not built (python -m oracle.build)")
        lib = ctypes.CDLL(LIB)
        P, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        PD = ctypes.POINTER(ctypes.c_double)
        lib.vvo_create.restype = P
        lib.vvo_create.argtypes = [I, I, D, D, D, I, I, D, PD, PD, PD, PD, D, D, I, D, I, I, I]
        lib.vvo_destroy.argtypes = [P]
        lib.vvo_step.restype = I
        lib.vvo_step.argtypes = [P, PD]
        lib.vvo_get_f.argtypes = [P, PD]
        lib.vvo_set_f.argtypes = [P, PD]
        lib.vvo_set_h.argtypes = [P, PD, PD, PD, PD]
        lib.vvo_residual.argtypes = [P, PD, PD, PD]
        lib.vvo_jacobian.argtypes = [P, PD, PD]
        lib.vvo_precond.argtypes = [P, PD, PD]
        lib.vvo_moments.argtypes = [P, PD, PD, PD, PD]
        lib.vvo_threads.restype = I
        lib.vvo_set_threads.argtypes = [I]
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class CRun:
    """C timestep on the natural-order field f (shape (nv, nx))."""

    def __init__(self, grid, f0_nat, bracket='arakawaj4', laplace='cfd4', charge=-1.0, mass=1.0,
                 ksp_rtol=1e-10, ksp_atol=1e-50, ksp_max_it=1000, snes_atol=1e-10, snes_max_iter=20,
                 restart=30, threads=0):
        lib = load()
        self.grid = grid
        self.snes_max_iter = int(snes_max_iter)
        self._v = np.ascontiguousarray(grid.v)
        e = O.kinetic_eigenvalues(grid.nx)
        self._eig = np.ascontiguousarray(np.stack([e.real, e.imag], axis=1).reshape(-1))
        self._f0 = np.ascontiguousarray(f0_nat, dtype=np.float64)
        self._h0 = np.ascontiguousarray(O.kinetic_hamiltonian(grid, mass))
        self.h = lib.vvo_create(grid.nx, grid.nv, grid.hx, grid.hv, grid.ht, BRACKETS[bracket],
                                4 if laplace == 'cfd4' else 2, charge, _p(self._v), _p(self._eig), _p(self._f0),
                                _p(self._h0), ksp_rtol, ksp_atol, ksp_max_it, snes_atol, snes_max_iter, restart,
                                int(threads))
        if not self.h:
            raise RuntimeError("vvo_create failed")
        self.lib = lib

    def step(self):
        """One timestep: Picard count, GMRES total, final residual and the
        residual after every Picard iteration (`residuals`)."""
        st = np.zeros(3 + self.snes_max_iter)
        rc = self.lib.vvo_step(self.h, _p(st))
        k = int(st[0])
        return dict(picard=k, gmres=int(st[1]), residual=float(st[2]), residuals=[float(x) for x in st[3:3 + k]],
                    rc=rc)

    def f(self):
        out = np.empty((self.grid.nv, self.grid.nx))
        self.lib.vvo_get_f(self.h, _p(out))
        return out

    def set_f(self, f_nat):
        """Replace the state (natural order (nv, nx)) and recompute its moments."""
        self._fs = np.ascontiguousarray(f_nat, dtype=np.float64)
        self.lib.vvo_set_f(self.h, _p(self._fs))

    def set_h(self, h1p, h1h, h2p, h2h):
        self._hs = [np.ascontiguousarray(a, dtype=np.float64) for a in (h1p, h1h, h2p, h2h)]
        self.lib.vvo_set_h(self.h, *[_p(a) for a in self._hs])

    def residual(self, f, fh):
        f, fh = np.ascontiguousarray(f), np.ascontiguousarray(fh)
        y = np.empty_like(f)
        self.lib.vvo_residual(self.h, _p(f), _p(fh), _p(y))
        return y

    def jacobian(self, u):
        u = np.ascontiguousarray(u)
        y = np.empty_like(u)
        self.lib.vvo_jacobian(self.h, _p(u), _p(y))
        return y

    def precond(self, x):
        x = np.ascontiguousarray(x)
        y = np.empty_like(x)
        self.lib.vvo_precond(self.h, _p(x), _p(y))
        return y

    def close(self):
        if self.h:
            self.lib.vvo_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def time_sample(w, steps=1, threads=0):
    """Time `steps` timesteps of workload w (dict from bench.workloads) after
    one untimed step, on `threads` OpenMP threads (0: all); returns the
    cpu_baseline record of bench.py.  The thread count in force before the
    call is restored."""
    lib = load()
    before = lib.vvo_threads()
    g = O.Grid(w["nx"], w["nv"], w["L"], w["vmax"], ht=w["ht"])
    f0 = O.normalise(O.initial_distribution(g, w["profile"]), g)
    try:
        run = CRun(g, f0.T, bracket=w["bracket"], threads=threads)
        cores = lib.vvo_threads()
        run.step()                                     # warm-up step (page faults, caches)
        t0 = time.perf_counter()
        stats = [run.step() for _ in range(steps)]
        dt = time.perf_counter() - t0
        run.close()
    finally:
        lib.vvo_set_threads(before)
    return {"value": round(steps / dt, 4), "unit": "timesteps/s", "cores": cores, "kind": "port",
            "sample": "%d timestep(s) of the same workload after 1 warm-up step (%.1f s); C/OpenMP "
                      "restatement of the timestep (oracle/c), %d thread(s); Picard/GMRES per step %s"
                      % (steps, dt, cores, [(s["picard"], s["gmres"]) for s in stats])}
