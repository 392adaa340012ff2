"""Many independent implicit Vlasov-Poisson runs on one GPU.

The reference runs one instance per process (run_nonlinear_matrixfree_split_mp.py);
this module runs a set of instances with uneven grids together:

* instances that share nx, nv, the v-grid, dt and the operator choice form a
  bucket; a bucket is one batched libvv1d context, so every kernel launch
  covers all of its instances (instance = second grid dimension);
* each bucket has its own HIP stream and host thread, so buckets of
  different shapes overlap on the GPU;
* each Newton (SNES ksponly) step of a bucket is one replay of a captured HIP
  graph (residual, P^-1, GMRES set-up and a full cycle of Arnoldi steps that
  return at once past an instance's convergence), when ``graph`` is set;
* launch shapes (FFT rows per workgroup, Gram-Schmidt vectors per thread) are
  chosen per bucket by timing the candidates on the bucket's own buffers
  (``autotune``).

Every instance follows the reference's Picard rules
(run_nonlinear_matrixfree_split_mp.py:130-168) on its own: instances whose
loop has stopped are masked out of the following linear solves, and a
residual that rose reverts that instance alone.
"""

import math
import threading
import time

import numpy as np
import torch

from .core import Grid, bump_on_tail, initial_distribution, kinetic_hamiltonian, landau, normalise
from .runner import ResidualTooLarge
from .solvers import Context


def make_instances(n=256, seed=0, vmax=10.0, ht=0.05, bracket='arakawaj4'):
    """The uneven-grid ensemble: nx from {64,128,256,512}, nv from
    {128,256,512,1024}, k ~ U[0.2, 0.6] with L = 2 pi / k, amplitude
    ~ U[0.005, 0.5], Landau or bump-on-tail profile with probability 1/2,
    all drawn from numpy.random.default_rng(seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(n):
        nx = int(rng.choice([64, 128, 256, 512]))
        nv = int(rng.choice([128, 256, 512, 1024]))
        k = float(rng.uniform(0.2, 0.6))
        amp = float(rng.uniform(0.005, 0.5))
        kind = 'landau' if rng.random() < 0.5 else 'bump_on_tail'
        prof = landau(k, amp) if kind == 'landau' else bump_on_tail(k, amp)
        out.append(dict(index=q, nx=nx, nv=nv, L=2 * math.pi / k, vmax=vmax, ht=ht, k=k, amp=amp,
                        profile_kind=kind, profile=prof, bracket=bracket))
    return out


def bucket_key(spec):
    return (spec['nx'], spec['nv'], spec['vmax'], spec['ht'], spec['bracket'])


class Bucket:
    """Instances of one shape, advanced together by one batched context."""

    def __init__(self, specs, ksp_rtol=1e-10, ksp_atol=1e-50, ksp_max_it=1000, snes_atol=1e-10,
                 snes_max_iter=20, restart=30, charge=-1.0, mass=1.0, laplace='cfd4', graph=True,
                 autotune=True, native=True):
        s0 = specs[0]
        self.specs = specs
        self.B = len(specs)
        self.grid = Grid(s0['nx'], s0['nv'], s0['L'], s0['vmax'], ht=s0['ht'])
        self.grids = [Grid(s['nx'], s['nv'], s['L'], s['vmax'], ht=s['ht']) for s in specs]
        for g in self.grids:
            assert g.nx == self.grid.nx and g.nv == self.grid.nv and np.array_equal(g.v, self.grid.v)
        self.ksp = (ksp_rtol, ksp_atol, 1e5, ksp_max_it)
        self.snes_atol, self.snes_max_iter = snes_atol, snes_max_iter
        self.stream = torch.cuda.Stream()
        with torch.cuda.stream(self.stream):
            ctx = self.ctx = Context(self.grid, bracket=s0['bracket'], preconditioner='tensor_kinetic',
                                     laplace=laplace, charge=charge, restart=restart,
                                     stream=self.stream.cuda_stream, hx=[g.hx for g in self.grids])
            self.h0 = torch.from_numpy(kinetic_hamiltonian(self.grid, mass)).cuda()
            self.fc, self.fh, self.fl, self.fy = (ctx.new_field() for _ in range(4))
            self.h1c, self.h1h, self.h2c, self.h2h = (ctx.new_x() for _ in range(4))
            self.n, self.phi, self.phih = ctx.new_x(), ctx.new_x(), ctx.new_x()
            f0 = np.stack([normalise(initial_distribution(g, s['profile']), g)
                           for g, s in zip(self.grids, specs)])
            self.fc.copy_(torch.from_numpy(f0))
            self.tuning = ctx.autotune() if autotune else ctx.get_tuning()
            ctx.moments(self.fc, self.n, self.phi, self.h1c, want_norm=False)
        if graph:
            ctx.set_graph(True)
        self.native = native
        self.itime = 0
        self.history = []

    def _update_previous(self):
        self.ctx.set_hamiltonian(self.h0, self.h1c, self.h1h, self.h2c, self.h2h)

    def _residual(self):
        # ||F(f)|| and the Poisson residual of the last moments call, one read
        fnorm, pnorm = self.ctx.residual_norms(self.fc, self.fh, self.fy)
        return fnorm + pnorm

    def step(self):
        """One timestep of every instance (the reference's run() body, per
        instance, with masks): in native code (vv_batched_timestep, no
        Python between the launches, so the buckets' host threads do not
        contend for the interpreter), or with ``native=False`` the same loop
        here."""
        if self.native:
            return self._step_native()
        ctx = self.ctx
        with torch.cuda.stream(self.stream):
            self.itime += 1
            self.fh.copy_(self.fc)
            self.h1h.copy_(self.h1c)
            self.h2h.copy_(self.h2c)
            self.phih.copy_(self.phi)
            self._update_previous()
            pred = self._residual()
            initial = pred.copy()
            active = np.ones(self.B, dtype=bool)
            picard = np.zeros(self.B, dtype=np.int32)
            gmres = np.zeros(self.B, dtype=np.int64)
            reverted = np.zeros(self.B, dtype=bool)
            history = []             # residual of every Picard iteration, per instance (NaN: not active)
            i = 0
            while active.any():
                i += 1
                ctx.set_active(active)
                self.fl.copy_(self.fc)
                res = ctx.snes_ksponly(self.fc, self.fh, *self.ksp)
                ctx.moments(self.fc, self.n, self.phi, self.h1c, want_norm=False)
                self._update_previous()
                new = self._residual()
                history.append(np.where(active, new, np.nan))
                revert = []
                for b in np.nonzero(active)[0]:
                    picard[b] += 1
                    gmres[b] += res[b].its
                    prev, pred[b] = pred[b], new[b]
                    if (pred[b] > prev and i > 1) or pred[b] < self.snes_atol or i >= self.snes_max_iter:
                        active[b] = False
                        if pred[b] > prev:
                            revert.append(b)
                if revert:
                    for b in revert:
                        self.fc[b].copy_(self.fl[b])
                        reverted[b] = True
                    ctx.moments(self.fc, self.n, self.phi, self.h1c, want_norm=False)
                    self._update_previous()
            ctx.set_active(None)
        rec = dict(step=self.itime, initial=initial, final=pred, picard=picard, gmres=gmres, reverted=reverted,
                   residuals=np.array(history).reshape(-1, self.B))
        self.history.append(rec)
        bad = np.nonzero(pred > 10.)[0]
        if bad.size:
            raise ResidualTooLarge("ERROR: Residual of nonlinear solver too large (instances %s)."
                                   % [self.specs[b]['index'] for b in bad])
        return rec

    def _step_native(self):
        fields = dict(f=self.fc, fh=self.fh, fl=self.fl, fy=self.fy, h0=self.h0, h1=self.h1c, h1h=self.h1h,
                      h2=self.h2c, h2h=self.h2h, n=self.n, phi=self.phi, phih=self.phih)
        with torch.cuda.stream(self.stream):
            self.itime += 1
            out, history, _ = self.ctx.batched_timestep(fields, *self.ksp, self.snes_atol, self.snes_max_iter)
        rec = dict(step=self.itime, initial=out["initial"], final=out["final"], picard=out["picard"].astype(np.int32),
                   gmres=out["gmres"].astype(np.int64), reverted=out["reverted"].astype(bool), residuals=history)
        self.history.append(rec)
        bad = np.nonzero(rec["final"] > 10.)[0]
        if bad.size:
            raise ResidualTooLarge("ERROR: Residual of nonlinear solver too large (instances %s)."
                                   % [self.specs[b]['index'] for b in bad])
        return rec

    def invariants(self):
        with torch.cuda.stream(self.stream):
            return self.ctx.invariants(self.fc, self.h0, self.h1c, self.h2c, self.phi)


class BatchRun:
    """All instances, bucketed; ``step()`` advances every instance by one
    timestep, buckets concurrently (one host thread and stream each)."""

    def __init__(self, specs, threads=True, **bucket_opts):
        groups = {}
        for s in specs:
            groups.setdefault(bucket_key(s), []).append(s)
        self.buckets = [Bucket(g, **bucket_opts) for _, g in sorted(groups.items())]
        self.threads = threads
        self.n_instances = len(specs)

    def step(self):
        if not self.threads or len(self.buckets) == 1:
            return [b.step() for b in self.buckets]
        out = [None] * len(self.buckets)
        err = []

        def work(q):
            try:
                out[q] = self.buckets[q].step()
            except BaseException as e:       # re-raised in the caller
                err.append(e)

        ts = [threading.Thread(target=work, args=(q,)) for q in range(len(self.buckets))]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        if err:
            raise err[0]
        return out

    def synchronize(self):
        for b in self.buckets:
            b.stream.synchronize()

    def state(self, index):
        """(f, phi) of one instance by its spec index, as numpy arrays."""
        for b in self.buckets:
            for q, s in enumerate(b.specs):
                if s['index'] == index:
                    b.stream.synchronize()
                    return b.fc[q].cpu().numpy(), b.phi[q].cpu().numpy()
        raise KeyError(index)


def sequential_config(spec, ksp_rtol=1e-10, snes_atol=1e-10, snes_max_iter=20, nt=1):
    """The Runscript configuration of one instance (for the sequential path)."""
    from .core import Config
    return Config({
        "grid": {"nx": spec['nx'], "nv": spec['nv'], "L": spec['L'], "vmax": spec['vmax'], "ht": spec['ht'],
                 "nt": nt},
        "solver": {"poisson_bracket": spec['bracket'], "laplace_operator": "CFD4",
                   "preconditioner_type": "tensor", "preconditioner_scheme": "tensor_kinetic",
                   "petsc_ksp_rtol": ksp_rtol, "petsc_ksp_atol": 1e-50, "petsc_ksp_max_iter": 1000,
                   "petsc_snes_atol": snes_atol, "petsc_snes_max_iter": snes_max_iter},
        "initial_data": {"charge": -1.0, "mass": 1.0},
    })


def run_sequential(specs, steps, **cfg_opts):
    """Each instance alone (single-instance contexts, one after the other).
    Returns (seconds spent in the timesteps, summed over the instances;
    {index: (f, phi)}); set-up is not timed, as for the batched run."""
    from .runner import Runscript
    elapsed = 0.0
    states = {}
    for s in specs:
        run = Runscript(sequential_config(s, nt=steps, **cfg_opts), profile=s['profile'])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            run.step()
        torch.cuda.synchronize()
        elapsed += time.perf_counter() - t0
        states[s['index']] = (run.fc.cpu().numpy(), run.pc_int.cpu().numpy())
        run.ctx.close()
    return elapsed, states
