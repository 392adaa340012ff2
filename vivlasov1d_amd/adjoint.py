"""Adjoint of the implicit timestep, behind torch.autograd.Function.

The reference has no derivatives.  A converged timestep f0 -> f1 of the
split implicit-midpoint scheme (run_nonlinear_matrixfree_split_mp.py:111-174)
satisfies

    G(f1, f0) = (f1 - f0)/dt + R {(f0 + f1)/2, H(f0, f1)} = 0,
    H = h0 + h1[f1]/2 + h1[f0]/2,

(R: rows 2..nv-3; h1[f] = the mean-free potential of the charge density of f,
Toolbox.pyx:38-98 with the Poisson solve of PETScPoissonSolver.pyx:58-74), so
by the implicit function theorem df1/df0 = -A^-1 A0 with A = dG/df1,
A0 = dG/df0.  The sensitivity of a function of f1 (lam1 = dE/df1) therefore
pulls back as

    A^T mu = lam1,   lam0 = -A0^T mu,

one transposed linear solve per timestep (vv_adjoint_step: GMRES on
P^-T A^T with the transposed kinetic preconditioner; the transposed bracket
comes from the skew-symmetry of the Arakawa stencils).  The forward states
are kept only every ``checkpoint`` steps and recomputed segment by segment in
the backward pass.

The result is the exact derivative of the converged discrete map; the Picard
loop's own stopping (snes_atol) bounds how closely the forward states satisfy
G = 0, and with it the agreement with finite differences.
"""

import numpy as np
import torch

from .runner import Runscript


class Stepper:
    """A Runscript whose state can be set from a tensor, with the pieces the
    adjoint needs (h1 of a state, the Hamiltonian of a step)."""

    def __init__(self, cfg, ksp_rtol_adj=1e-12, checkpoint=20):
        self.run = Runscript(cfg, f0=np.zeros((cfg['grid']['nv'], cfg['grid']['nx'])), normalise_f0=False)
        self.ctx = self.run.ctx
        self.grid = self.run.grid
        self.checkpoint = checkpoint
        self.ksp_adj = (ksp_rtol_adj, 1e-300, 1e5, 1000)
        ctx = self.ctx
        self._n, self._phi, self._h1a, self._h1b = ctx.new_x(), ctx.new_x(), ctx.new_x(), ctx.new_x()
        self._zx = ctx.new_x()
        self.adjoint_its = []

    def advance(self, f, nsteps):
        """nsteps implicit timesteps from f; returns the list of states
        [f, f_1, ..., f_nsteps] (CUDA tensors)."""
        self.run.set_state(f)
        out = [self.run.fc.clone()]
        for _ in range(nsteps):
            self.run.step()
            out.append(self.run.fc.clone())
        return out

    def h1(self, f, out):
        self.ctx.moments(f, self._n, self._phi, out, want_norm=False)

    def pullback(self, lam1, f0, f1):
        """lam0 = -A0^T A^-T lam1 for the step f0 -> f1."""
        ctx = self.ctx
        self.h1(f1, self._h1a)
        self.h1(f0, self._h1b)
        ctx.set_hamiltonian(self.run.h0, self._h1a, self._h1b, self._zx, self._zx)
        fbar = 0.5 * f0 + 0.5 * f1
        lam0 = torch.empty_like(lam1)
        mu = torch.empty_like(lam1)
        res = ctx.adjoint_step(lam1.contiguous(), lam0, mu, *self.ksp_adj, fbar=fbar)
        if res.reason <= 0:
            raise RuntimeError("adjoint linear solve did not converge (reason %d)" % res.reason)
        self.adjoint_its.append(res.its)
        return lam0


class ImplicitTimesteps(torch.autograd.Function):
    """f_K = Phi^K(f_0) for the implicit timestep Phi; backward by the
    discrete adjoint with checkpointed recomputation."""

    @staticmethod
    def forward(ctx, f0, stepper, nsteps):
        cp = max(1, int(stepper.checkpoint))
        checkpoints = {0: f0.detach().clone()}
        f = f0.detach()
        done = 0
        while done < nsteps:
            seg = min(cp, nsteps - done)
            states = stepper.advance(f, seg)
            f = states[-1]
            done += seg
            checkpoints[done] = f.clone()
        ctx.stepper, ctx.nsteps, ctx.cp, ctx.checkpoints = stepper, nsteps, cp, checkpoints
        return f.clone()

    @staticmethod
    def backward(ctx, grad):
        st = ctx.stepper
        lam = grad.contiguous().clone()
        marks = sorted(ctx.checkpoints)
        for q in range(len(marks) - 1, 0, -1):
            a, b = marks[q - 1], marks[q]
            states = st.advance(ctx.checkpoints[a], b - a)     # recompute the segment
            for n in range(b - a - 1, -1, -1):
                lam = st.pullback(lam, states[n], states[n + 1])
        return lam, None, None


class FieldEnergy(torch.autograd.Function):
    """E = 1/2 q sum_i (phi_{i+1} - phi_i)^2 / hx of the potential of f: the
    field energy of potential.pyx:48-59 (calculate_energy), charge factor
    included (negative for q = -1)."""

    @staticmethod
    def forward(ctx, f, stepper):
        c = stepper.ctx
        n, phi, h1 = c.new_x(), c.new_x(), c.new_x()
        c.moments(f.detach().contiguous(), n, phi, h1, want_norm=False)
        d = torch.roll(phi, -1) - phi
        ctx.stepper, ctx.phi = stepper, phi
        return 0.5 * stepper.run.charge * (d * d).sum() / stepper.grid.hx

    @staticmethod
    def backward(ctx, grad):
        st, phi = ctx.stepper, ctx.phi
        g = st.grid
        # dE/dphi = q D^T D phi / hx (D: periodic forward difference)
        dphi = st.run.charge * ((phi - torch.roll(phi, 1)) - (torch.roll(phi, -1) - phi)) / g.hx
        # dE/dn = (-q) (I - mean) C dE/dphi: the Poisson solve of the moments
        # pipeline applied to dE/dphi (mean-free already); dn/df = hv sum_j
        w = st.ctx.new_x()
        st.ctx.poisson_solve(dphi.contiguous(), w)
        lam = (g.hv * w).unsqueeze(0).expand(g.nv, g.nx).contiguous()
        return grad * lam, None


def perturbation_basis(grid, k, nmodes=16):
    """cos(m k x) M(v) on rows 2..nv-3, m = 1..nmodes, as a (nmodes, nv, nx)
    float64 array (M the unit Maxwellian)."""
    x = grid.x
    v = grid.v
    m = np.exp(-0.5 * v ** 2) / np.sqrt(2 * np.pi)
    out = np.zeros((nmodes, grid.nv, grid.nx))
    for q in range(nmodes):
        out[q] = np.outer(m, np.cos((q + 1) * k * x))
    out[:, :2] = 0.0
    out[:, -2:] = 0.0
    return out
