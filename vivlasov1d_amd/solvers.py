"""Solver objects with the reference's class and callback interface, backed
by libvv1d.so on the GPU.

Reference counterparts (paths in the viVlasov1D tree):

* ``PETScVlasovSolver``      vlasov/solvers/vlasov/PETScVlasovSolver.pyx:19-205
                             + PETScNLVlasovMP.pyx:15-58
* ``TensorProductPreconditioner``  vlasov/solvers/preconditioner/TensorProduct.pyx:66-86
* ``PoissonBracket``, ``TimeDerivative``  vlasov/solvers/components/*.pyx factories
* ``PETScPoissonSolver``     vlasov/solvers/poisson/PETScPoissonSolver.pyx,
                             PETScPoissonCFD4.pyx, PETScPoissonCFD2.pyx
* ``Toolbox``                vlasov/toolbox/Toolbox.pyx:38-224
* ``KSP``, ``SNES``          the PETSc objects configured at
                             run_nonlinear_matrixfree_split_mp.py:42-80

Vectors are torch float64 tensors on the GPU.  A 2D field is C-contiguous
with shape (nv, nx): element (i, j) at i + nx*j, the reference's PETSc
natural order.  x-vectors have shape (nx,), v-vectors (nv,).  The
Hamiltonian pieces are passed by their 1D factors: H0 as the (nv,) kinetic
profile, H1/H2 as (nx,) potentials -- the only form the reference ever
builds (Toolbox.pyx:59, :187).
"""

import ctypes
import warnings

import numpy as np
import torch

from . import _lib
from .core import Grid, kinetic_eigenvalues

DOUBLE = torch.float64


def _dev_ptr(t, numel=None, name="vector", dtype=DOUBLE):
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch tensor" % name)
    if t.dtype != dtype or not t.is_cuda or not t.is_contiguous():
        raise ValueError("%s must be a contiguous %s CUDA tensor" % (name, str(dtype).replace("torch.", "")))
    if numel is not None and t.numel() != numel:
        raise ValueError("%s has %d elements, expected %d" % (name, t.numel(), numel))
    return ctypes.c_void_p(t.data_ptr())


class Context:
    """One libvv1d context: grid, operator choice and device scratch.

    With ``hx`` (a sequence of x-spacings) the context is batched: it holds
    len(hx) instances that share the grid's nx, nv, v-grid and dt, each with
    its own x-period.  Fields are then (B, nv, nx) tensors, x-vectors (B, nx),
    and the scalar-returning methods return one value per instance."""

    def __init__(self, grid, bracket='arakawaj4', preconditioner='tensor_kinetic', laplace='cfd4',
                 charge=-1.0, restart=30, stream=None, hx=None):
        lib = _lib.load()
        if bracket not in _lib.BRACKETS:
            raise _lib.VVError("Poisson bracket %r is not provided (arakawaj1/2/4)" % (bracket,))
        if preconditioner not in _lib.PRECONDITIONERS:
            raise _lib.VVError("preconditioner %r is not provided (tensor_kinetic, tensor_potential or none)"
                               % (preconditioner,))
        if laplace not in _lib.LAPLACE:
            raise _lib.VVError("laplace operator %r is not provided (cfd2/cfd4)" % (laplace,))
        self.grid = grid
        self.bracket, self.preconditioner, self.laplace = bracket, preconditioner, laplace
        self.charge = float(charge)
        self._v = np.ascontiguousarray(grid.v, dtype=np.float64)
        # TensorProduct.compute_eigenvalues of the transformed direction: x for
        # tensor_kinetic, v for tensor_potential
        self.potential = _lib.PRECONDITIONERS[preconditioner] == 2
        eig = kinetic_eigenvalues(grid.nv if self.potential else grid.nx)
        self._eig = np.ascontiguousarray(np.stack([eig.real, eig.imag], axis=1).reshape(-1))
        p = _lib.VVParams()
        p.nx, p.nv = grid.nx, grid.nv
        p.hx, p.hv, p.ht = grid.hx, grid.hv, grid.ht
        p.charge = self.charge
        p.bracket = _lib.BRACKETS[bracket]
        p.precond = _lib.PRECONDITIONERS[preconditioner]
        p.laplace = _lib.LAPLACE[laplace]
        p.restart = int(restart)
        p.v = self._v.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        p.eigen = self._eig.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        self._ptr = ctypes.c_void_p()
        if hx is None:
            self.B = 1
            self.batched = False
            ret = lib.vv_ctx_create(ctypes.byref(p), ctypes.c_void_p(stream), ctypes.byref(self._ptr))
        else:
            self._hx = np.ascontiguousarray(hx, dtype=np.float64)
            self.B = int(self._hx.size)
            self.batched = True
            p.hx = float(self._hx[0])
            ret = lib.vv_ctx_create_batched(ctypes.byref(p), ctypes.c_int32(self.B),
                                            self._hx.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                            ctypes.c_void_p(stream), ctypes.byref(self._ptr))
        _lib.check(ret, None)
        self.lib = lib
        self.stream = stream
        self.N = grid.nx * grid.nv
        self.dissipation = None

    @property
    def handle(self):
        return self._ptr

    def scratch_bytes(self):
        """Device bytes the context owns (vv_scratch_bytes)."""
        return int(self.lib.vv_scratch_bytes(self._ptr))

    def call(self, name, *args):
        return _lib.check(getattr(self.lib, name)(self._ptr, *args), self._ptr)

    def close(self):
        if getattr(self, "_ptr", None) and self._ptr.value:
            self.lib.vv_ctx_destroy(self._ptr)
            self._ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # fields ----------------------------------------------------------------
    def F(self, t, name="field"):
        return _dev_ptr(t, self.N * self.B, name)

    def X(self, t, name="x-vector"):
        return _dev_ptr(t, self.grid.nx * self.B, name)

    def V(self, t, name="v-vector"):
        return _dev_ptr(t, self.grid.nv, name)

    def new_field(self):
        shape = (self.B, self.grid.nv, self.grid.nx) if self.batched else (self.grid.nv, self.grid.nx)
        return torch.zeros(shape, dtype=DOUBLE, device='cuda')

    def new_x(self):
        shape = (self.B, self.grid.nx) if self.batched else (self.grid.nx,)
        return torch.zeros(shape, dtype=DOUBLE, device='cuda')

    def _per_instance(self, arr):
        return np.array(arr[:self.B], dtype=np.float64) if self.batched else float(arr[0])

    # batching, graphs, launch shapes ------------------------------------------
    def set_active(self, active=None):
        if active is None:
            self.call("vv_set_active", None)
        else:
            a = (ctypes.c_int32 * self.B)(*[int(bool(v)) for v in active])
            self.call("vv_set_active", a)

    def set_stream(self, handle):
        """Enqueue on another HIP stream (a hipStream_t as an int; 0 = null stream)."""
        self.call("vv_set_stream", ctypes.c_void_p(handle))
        self.stream = handle

    def set_graph(self, enable=True):
        self.call("vv_set_graph", ctypes.c_int(1 if enable else 0))

    def set_tuning(self, fft_rows=0, cgs_ept=0):
        self.call("vv_set_tuning", ctypes.c_int32(fft_rows), ctypes.c_int32(cgs_ept))

    def get_tuning(self):
        r, e = ctypes.c_int32(), ctypes.c_int32()
        self.call("vv_get_tuning", ctypes.byref(r), ctypes.byref(e))
        return r.value, e.value

    def set_fusion(self, enable=True):
        """Fused kernels in the Arnoldi step (default on): irFFT + Gram-Schmidt
        dots, and J v + rFFT; results identical up to the sign of exact zeros."""
        self.call("vv_set_fusion", ctypes.c_int(1 if enable else 0))

    def autotune(self):
        """Time the launch-shape candidates; returns (fft_rows, cgs_ept) kept."""
        r, e = ctypes.c_int32(), ctypes.c_int32()
        self.call("vv_autotune", ctypes.byref(r), ctypes.byref(e))
        return r.value, e.value

    # operators ---------------------------------------------------------------
    def set_hamiltonian(self, h0, h1p, h1h, h2p, h2h):
        self.call("vv_set_hamiltonian", self.V(h0, "H0"), self.X(h1p, "H1p"), self.X(h1h, "H1h"),
                  self.X(h2p, "H2p"), self.X(h2h, "H2h"))

    def residual(self, f, fh, y):
        self.call("vv_residual", self.F(f), self.F(fh), self.F(y))

    def jacobian_apply(self, u, y):
        self.call("vv_jacobian_apply", self.F(u), self.F(y))

    def bracket_add(self, x, y, factor):
        self.call("vv_bracket_add", self.F(x), self.F(y), ctypes.c_double(factor))

    def precond_update(self, phi):
        """update_matrices: the potential preconditioner's bands from phi (an
        x-vector per instance); a no-op for tensor_kinetic."""
        self.call("vv_precond_update", self.X(phi))

    def precond(self, x, y):
        self.call("vv_precond_apply", self.F(x), self.F(y))

    def function(self, f, fh, y):
        self.call("vv_function", self.F(f), self.F(fh), self.F(y))

    def jacobian(self, u, y):
        self.call("vv_jacobian", self.F(u), self.F(y))

    def function_norm(self, f, fh, y):
        out = (ctypes.c_double * self.B)()
        self.call("vv_function_norm", self.F(f), self.F(fh), self.F(y), out)
        return self._per_instance(out)

    def residual_norms(self, f, fh, y):
        """(||F(f)||_2, Poisson residual norm of the last moments call) in one read."""
        fo, po = (ctypes.c_double * self.B)(), (ctypes.c_double * self.B)()
        self.call("vv_residual_norms", self.F(f), self.F(fh), self.F(y), fo, po)
        return self._per_instance(fo), self._per_instance(po)

    def density(self, f, n):
        self.call("vv_density", self.F(f), self.X(n))

    def velocity_moments(self, f, u, e):
        """u = hv sum_j v_j f_ij, e = hv sum_j v_j^2 f_ij (Toolbox.pyx:103-148)."""
        self.call("vv_velocity_moments", self.F(f), self.X(u), self.X(e))

    # dissipation (CollisionOperator.pyx, DoubleBracket.pyx) ------------------
    def set_dissipation(self, kind, coll_freq=0., coll_diff=1., coll_drag=1.):
        """kind: 'collt' | 'colle' | 'ffh' | 'fhh' | None; nothing is added
        when coll_freq == 0 (config.py:148-163)."""
        if kind not in _lib.DISSIPATION:
            raise _lib.VVError("dissipation %r is not provided (collt, colle, ffh, fhh)" % (kind,))
        d = _lib.Dissipation()
        d.kind = _lib.DISSIPATION[kind]
        d.coll_freq, d.coll_diff, d.coll_drag = float(coll_freq), float(coll_diff), float(coll_drag)
        self.call("vv_set_dissipation", ctypes.byref(d))
        self.dissipation = kind if (d.kind and coll_freq != 0.) else None

    def set_moments(self, np_, up, ep, ap, nh, uh, eh, ah):
        """The moment x-vectors the collision operators read; the tensors must
        stay alive (the context keeps their addresses, PETScVlasovSolver.pyx:123-132)."""
        ptr = lambda t: None if t is None else self.X(t)
        self._moments = (np_, up, ep, ap, nh, uh, eh, ah)
        self.call("vv_set_moments", *[ptr(t) for t in self._moments])

    def collision_factor(self, n, u, e, a):
        """a = n n / (n e - u u) (Toolbox.pyx:153-165)."""
        self.call("vv_collision_factor", self.X(n), self.X(u), self.X(e), self.X(a))

    def collision_add(self, f, y, n, u, e, a, factor):
        self.call("vv_collision_add", self.F(f), self.F(y), self.X(n), self.X(u),
                  None if e is None else self.X(e), self.X(a), ctypes.c_double(factor))

    def double_bracket_function(self, fave, y, factor):
        self.call("vv_double_bracket_function", self.F(fave), self.F(y), ctypes.c_double(factor))

    def double_bracket_jacobian(self, u, fave, y, factor):
        self.call("vv_double_bracket_jacobian", self.F(u), self.F(fave), self.F(y), ctypes.c_double(factor))

    # explicit steppers of the initial guess (run_base.py:786-927) ----------
    def symplectic_kinetic(self, x, h0, factor, niter=1):
        """x += factor*ht/niter * (-{x, H0})  (PETScArakawaSymplectic.pyx:62-77)."""
        self.call("vv_symplectic_kinetic", self.F(x), self.V(h0), ctypes.c_double(factor), ctypes.c_int32(niter))

    def symplectic_potential(self, x, h1, h2, factor, niter=1):
        """x += factor*ht/niter * (-{x, H1 + H2})  (PETScArakawaSymplectic.pyx:80-97)."""
        self.call("vv_symplectic_potential", self.F(x), self.X(h1), self.X(h2), ctypes.c_double(factor),
                  ctypes.c_int32(niter))

    def rk4(self, x, h0, h1, h2, rule='rk4_16', niter=1):
        """One RK4 step of dt/niter with H = H0 + H1 + H2 (PETScArakawaRungeKutta.pyx:20-98)."""
        self.call("vv_rk4", self.F(x), self.V(h0), self.X(h1), self.X(h2), ctypes.c_int32(_lib.RK4_RULES[rule]),
                  ctypes.c_int32(niter))

    def poisson_solve(self, n, phi):
        self.call("vv_poisson_solve", self.X(n), self.X(phi))

    def poisson_matrix_mult(self, x, y):
        self.call("vv_poisson_matrix_mult", self.X(x), self.X(y))

    def poisson_function_mult(self, x, n, y):
        self.call("vv_poisson_function_mult", self.X(x), self.X(n), self.X(y))

    def potential_to_hamiltonian(self, phi, h1):
        self.call("vv_potential_to_hamiltonian", self.X(phi), self.X(h1))

    def moments(self, f, n, phi, h1, want_norm=True):
        out = (ctypes.c_double * self.B)()
        self.call("vv_moments", self.F(f), self.X(n), self.X(phi), self.X(h1), out if want_norm else None)
        return self._per_instance(out) if want_norm else None

    def moments_noalloc(self, f, n, phi, h1, pnorm=None):
        """vv_moments_noalloc: the moments pipeline without allocation; the
        Poisson residual norms go to the device tensor pnorm (batch doubles)."""
        self.call("vv_moments_noalloc", self.F(f), self.X(n), self.X(phi), self.X(h1),
                  None if pnorm is None else _dev_ptr(pnorm, self.B, "pnorm"))

    def norm2(self, x):
        out = ctypes.c_double()
        self.call("vv_norm2", _dev_ptr(x), ctypes.c_int64(x.numel()), ctypes.byref(out))
        return out.value

    def dot(self, x, y):
        out = ctypes.c_double()
        if x.numel() != y.numel():
            raise ValueError("dot: size mismatch")
        self.call("vv_dot", _dev_ptr(x), _dev_ptr(y), ctypes.c_int64(x.numel()), ctypes.byref(out))
        return out.value

    def axpy(self, alpha, x, y):
        if x.numel() != y.numel():
            raise ValueError("axpy: size mismatch")
        self.call("vv_axpy", ctypes.c_double(alpha), _dev_ptr(x), _dev_ptr(y), ctypes.c_int64(x.numel()))

    def ksp_solve(self, b, x, rtol, atol, dtol, max_it):
        """KspResult (one instance) or a list of them (batched)."""
        kp = _lib.KspParams(rtol, atol, dtol, int(max_it))
        res = (_lib.KspResult * self.B)()
        self.call("vv_ksp_solve", self.F(b), self.F(x), ctypes.byref(kp), res)
        return list(res) if self.batched else res[0]

    def snes_ksponly(self, f, fh, rtol, atol, dtol, max_it):
        kp = _lib.KspParams(rtol, atol, dtol, int(max_it))
        res = (_lib.KspResult * self.B)()
        self.call("vv_snes_ksponly", self.F(f), self.F(fh), ctypes.byref(kp), res)
        return list(res) if self.batched else res[0]

    def batched_timestep(self, fields, rtol, atol, dtol, max_it, snes_atol, snes_max_iter):
        """vv_batched_timestep: one timestep of every instance, the Picard loop
        in native code.  fields: dict of the device tensors f, fh, fl, fy
        (fields), h0 (v-vector), h1, h1h, h2, h2h, n, phi, phih (x-vectors).
        Returns (per-instance records as a dict of numpy arrays, residual
        history rows x B with NaN where inactive, iterations)."""
        bf = _lib.BatchFields()
        for k in ("f", "fh", "fl", "fy"):
            setattr(bf, k, self.F(fields[k], k).value)
        bf.h0 = self.V(fields["h0"], "h0").value
        for k in ("h1", "h1h", "h2", "h2h", "n", "phi", "phih"):
            setattr(bf, k, self.X(fields[k], k).value)
        kp = _lib.KspParams(rtol, atol, dtol, int(max_it))
        rec = (_lib.InstanceStep * self.B)()
        rows = int(snes_max_iter)
        hist = (ctypes.c_double * (rows * self.B))()
        its = ctypes.c_int32()
        self.call("vv_batched_timestep", ctypes.byref(bf), ctypes.byref(kp), ctypes.c_double(snes_atol),
                  ctypes.c_int32(rows), rec, hist, ctypes.c_int32(rows), ctypes.byref(its))
        n = its.value
        out = {k: np.array([getattr(r, k) for r in rec]) for k in ("picard", "reverted", "gmres", "initial", "final")}
        history = np.array(hist[:n * self.B], dtype=np.float64).reshape(n, self.B)
        return out, history, n

    # mixed precision ----------------------------------------------------------
    def F32(self, t, name="field"):
        return _dev_ptr(t, self.N * self.B, name, torch.float32)

    def jacobian_apply32(self, u, y):
        self.call("vv_jacobian_apply32", self.F32(u), self.F32(y))

    def precond32(self, x, y):
        self.call("vv_precond_apply32", self.F32(x), self.F32(y))

    def ksp_solve_mixed(self, b, x, rtol, atol, dtol, max_it, inner_rtol=1e-5, max_cycles=50, fp64_check=True):
        """FP32 GMRES cycles; returns (KspResult, cycles).  fp64_check: refine
        on the true FP64 residual after every cycle; False: one FP32 cycle with
        the given tolerances decides (for an outer loop that refines)."""
        kp = _lib.KspParams(rtol, atol, dtol, int(max_it))
        mp = _lib.MixedParams(inner_rtol, int(max_cycles), 1 if fp64_check else 0)
        res = _lib.KspResult()
        cyc = ctypes.c_int32()
        self.call("vv_ksp_solve_mixed", self.F(b), self.F(x), ctypes.byref(kp), ctypes.byref(mp),
                  ctypes.byref(res), ctypes.byref(cyc))
        return res, cyc.value

    def snes_ksponly_mixed(self, f, fh, rtol, atol, dtol, max_it, inner_rtol=1e-5, max_cycles=50, fp64_check=True):
        kp = _lib.KspParams(rtol, atol, dtol, int(max_it))
        mp = _lib.MixedParams(inner_rtol, int(max_cycles), 1 if fp64_check else 0)
        res = _lib.KspResult()
        cyc = ctypes.c_int32()
        self.call("vv_snes_ksponly_mixed", self.F(f), self.F(fh), ctypes.byref(kp), ctypes.byref(mp),
                  ctypes.byref(res), ctypes.byref(cyc))
        return res, cyc.value

    # linearisation and adjoint ------------------------------------------------
    def linear_apply(self, u, y, which, fbar=None):
        """y = A u, A0 u, A^T u or A0^T u (which = 'A', 'A0', 'AT', 'A0T') of
        the step linearised about the H_ave set last and f_bar (None: no
        field coupling)."""
        self.call("vv_linear_apply", self.F(u), self.F(fbar) if fbar is not None else None, self.F(y),
                  ctypes.c_int32(_lib.LINEAR[which]))

    def precond_t(self, x, y):
        self.call("vv_precond_apply_t", self.F(x), self.F(y))

    def adjoint_solve(self, lam, mu, rtol, atol, dtol, max_it, fbar=None):
        kp = _lib.KspParams(rtol, atol, dtol, int(max_it))
        res = (_lib.KspResult * self.B)()
        self.call("vv_adjoint_solve", self.F(lam), self.F(fbar) if fbar is not None else None, self.F(mu),
                  ctypes.byref(kp), res)
        return list(res) if self.batched else res[0]

    def adjoint_step(self, lam1, lam0, mu, rtol, atol, dtol, max_it, fbar=None):
        kp = _lib.KspParams(rtol, atol, dtol, int(max_it))
        res = (_lib.KspResult * self.B)()
        self.call("vv_adjoint_step", self.F(lam1), self.F(fbar) if fbar is not None else None, self.F(lam0),
                  self.F(mu), ctypes.byref(kp), res)
        return list(res) if self.batched else res[0]

    def invariants(self, f, h0, h1, h2, phi):
        out = (ctypes.c_double * (6 * self.B))()
        self.call("vv_invariants", self.F(f), self.V(h0), self.X(h1), self.X(h2), self.X(phi), out)
        keys = ("mass", "l2", "energy", "ekin", "epot", "efield")
        vals = list(out)
        res = [dict(zip(keys, vals[6 * b:6 * b + 6])) for b in range(self.B)]
        return res if self.batched else res[0]

    PROFILE_KERNELS = ("jacobian", "residual", "rfft", "tridiag", "irfft", "mdot", "maxpy_norm",
                       "jacobian32", "rfft32", "tridiag32", "irfft32", "mdot32", "maxpy_norm32",
                       "irfft_mdot", "irfft_mdot32", "jac_rfft", "jac_rfft32")

    def profile(self, kernels=None, every=1):
        """Time the named kernels (all when None; () switches timing off),
        one launch in every `every` of each."""
        self.call("vv_profile_sample", ctypes.c_int(int(every)))
        if kernels is None:
            mask = -1
        else:
            mask = 0
            for k in kernels:
                mask |= 1 << self.PROFILE_KERNELS.index(k)
        self.call("vv_profile_enable", ctypes.c_int(mask))

    def profile_read(self, kernel):
        """(total ms, launches, algorithmic bytes) of one kernel since the last read."""
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        by = ctypes.c_double()
        self.call("vv_profile_read", kernel.encode(), ctypes.byref(ms), ctypes.byref(n), ctypes.byref(by))
        return ms.value, n.value, by.value

    def synchronize(self):
        self.call("vv_synchronize")


# ---------------------------------------------------------------------------
# component factories (names and keys of the reference)
# ---------------------------------------------------------------------------
class PoissonBracket:
    """PoissonBracket.create (PoissonBracket.pyx:37-50): 'arakawaj1' | 'arakawaj2' | 'arakawaj4'."""

    KINDS = ('arakawaj1', 'arakawaj2', 'arakawaj4')

    def __init__(self, kind, context):
        self.kind = kind
        self.ctx = context

    @staticmethod
    def create(type, da1, grid, context=None):
        if type not in PoissonBracket.KINDS:
            if type == 'simpson':
                raise _lib.VVError("the Simpson bracket is not provided by this library")
            return None
        ctx = context if context is not None else Context(grid, bracket=type, preconditioner='')
        return PoissonBracket(type, ctx)

    def function(self, F, H, Y, factor):
        """Y += factor {F, H} (PoissonBracket.pyx:54-63).  H must be the
        context's H_ave (set through update_previous / set_hamiltonian);
        H is accepted for signature parity."""
        self.ctx.bracket_add(F, Y, factor)


class CollisionOperator:
    """CollisionOperator.create (CollisionOperator.pyx:43-57): 'collt' |
    'colle' build the operator on the context, anything else the no-op base."""

    def __init__(self, kind, context, coll_freq=0., coll_diff=1., coll_drag=1.):
        self.kind = kind
        self.ctx = context
        self.coll_freq, self.coll_diff, self.coll_drag = coll_freq, coll_diff, coll_drag

    @staticmethod
    def create(type, da1, grid, coll_freq=0., coll_diff=1., coll_drag=1., context=None):
        kind = type if type in ('collt', 'colle') else None
        ctx = context if context is not None else Context(grid, preconditioner='')
        if kind is not None:
            ctx.set_dissipation(kind, coll_freq, coll_diff, coll_drag)
        return CollisionOperator(kind, ctx, coll_freq, coll_diff, coll_drag)

    def function(self, F, Y, N, U, E, A, factor):
        """Y += C(F) with u = U/N (:72-110, :164-202); a no-op for the base
        class and when coll_freq or factor is 0."""
        if self.kind is None or self.coll_freq == 0. or factor == 0.:
            return
        self.ctx.collision_add(F, Y, N, U, E, A, factor)

    def jacobian(self, J, N, U, E, A, factor):
        raise _lib.VVError("the assembled Jacobian (direct-solver path) is not provided")


class DoubleBracket:
    """DoubleBracket.create (DoubleBracket.pyx:43-56): 'ffh' | 'fhh' build the
    term on the context, anything else the no-op base.  H must be the
    context's H_ave (set through update_previous / set_hamiltonian); the Have
    arguments are accepted for signature parity."""

    def __init__(self, kind, context, coll_freq=0.):
        self.kind = kind
        self.ctx = context
        self.coll_freq = coll_freq

    @staticmethod
    def create(type, da1, grid, poisson_bracket, coll_freq=0., context=None):
        kind = type if type in ('ffh', 'fhh') else None
        ctx = context if context is not None else poisson_bracket.ctx
        if kind is not None:
            ctx.set_dissipation(kind, coll_freq)
        return DoubleBracket(kind, ctx, coll_freq)

    def function(self, Fave, Have, Y, factor):
        if self.kind is not None:
            self.ctx.double_bracket_function(Fave, Y, factor)

    def jacobian(self, F, Fave, Have, Y, factor):
        if self.kind is not None:
            self.ctx.double_bracket_jacobian(F, Fave, Y, factor)


class PETScArakawaSymplectic:
    """PETScArakawaSymplectic (solvers/explicit/PETScArakawaSymplectic.pyx:17-97)
    on the GPU.  H0 is the kinetic Hamiltonian's v-vector, H1 and H2 the
    x-vectors of the potentials (kept by reference, as the reference keeps
    its Vecs); niter divides the time step."""

    def __init__(self, config, da1, grid, H0, H1, H2, niter=1, context=None):
        self.grid, self.niter = grid, int(niter)
        self.H0, self.H1, self.H2 = H0, H1, H2
        self.ctx = context if context is not None else Context(grid, bracket=config.get_poisson_bracket(),
                                                               preconditioner='')

    def kinetic(self, X, factor=0.5):
        self.ctx.symplectic_kinetic(X, self.H0, factor, self.niter)

    def potential(self, X, factor=1.0):
        self.ctx.symplectic_potential(X, self.H1, self.H2, factor, self.niter)


class PETScArakawaRungeKutta:
    """PETScArakawaRungeKutta (solvers/explicit/PETScArakawaRungeKutta.pyx:15-98)
    on the GPU: one RK4 step of dt/niter with H = H0 + H1 + H2."""

    def __init__(self, config, da1, grid, H0, H1, H2, niter=1, context=None):
        self.grid, self.niter = grid, int(niter)
        self.H0, self.H1, self.H2 = H0, H1, H2
        self.ctx = context if context is not None else Context(grid, bracket=config.get_poisson_bracket(),
                                                               preconditioner='')

    def rk4_16(self, X, H1=None):
        self.ctx.rk4(X, self.H0, self.H1, self.H2, 'rk4_16', self.niter)

    def rk4_18(self, X, H1=None):
        self.ctx.rk4(X, self.H0, self.H1, self.H2, 'rk4_18', self.niter)


class TimeDerivative:
    """TimeDerivative.create (TimeDerivative.pyx:36-53): only 'point'."""

    def __init__(self, grid):
        self.grid = grid

    @staticmethod
    def create(type, da1, grid):
        if type == 'point':
            return TimeDerivative(grid)
        if type in ('midpoint', 'simpson', 'arakawa_J1', 'arakawa_J2', 'arakawa_J4'):
            raise _lib.VVError("averaged time derivative %r is not provided by this library" % type)
        return None

    def function(self, F, Y):
        """Y += F / dt (TimeDerivative.pyx:56-57)."""
        Y.add_(F, alpha=self.grid.ht_inv)


class TensorProductPreconditioner:
    """TensorProductPreconditioner.create (TensorProduct.pyx:66-86): the
    kinetic (FFT along x, tridiagonal along v) and the potential (FFT along v,
    tridiagonal along x with phi-dependent bands) forms; the _scipy schemes
    state the same operators.  ``phi`` is the potential the potential form's
    bands follow (refreshed by update_matrices)."""

    def __init__(self, context, phi=None):
        self.ctx = context
        self.phi = phi

    @staticmethod
    def create(type, da1, daphi, grid, phi, context=None):
        if type in ('tensor_kinetic', 'tensor_kinetic_scipy', 'tensor_potential', 'tensor_potential_scipy'):
            ctx = context if context is not None else Context(grid, preconditioner=type)
            pc = TensorProductPreconditioner(ctx, phi)
            pc.update_matrices()
            return pc
        return None

    def tensorProduct(self, X, Y):
        """Y = P^-1 X (TensorProductKinetic.pyx:146-160, TensorProductPotential.pyx:157-173)."""
        self.ctx.precond(X, Y)

    def update_matrices(self):
        """The potential form's bands from phi (PotentialFast.pyx:205-264); the
        kinetic form depends on the grid only (KineticFast.pyx:244-245)."""
        if self.ctx.potential and self.phi is not None:
            self.ctx.precond_update(self.phi)


def effective_preconditioner(scheme):
    """The preconditioner a configured scheme gives: TensorProductPreconditioner.create
    (TensorProduct.pyx:66-86) builds the four tensor schemes and returns None
    for every other name ('TensorFast' in examples/twostream_bgk*.cfg,
    'tensor_diagonal', ...), so such a run goes on without a preconditioner.
    The same happens here, with a warning."""
    if scheme in _lib.PRECONDITIONERS:
        return scheme
    warnings.warn("preconditioner scheme %r is not tensor_kinetic[_scipy] or tensor_potential[_scipy]: "
                  "running without a preconditioner, as the reference does" % (scheme,))
    return ''


class PETScVlasovSolver:
    """PETScVlasovSolverBase + PETScNLVlasovMP on the GPU.

    Constructor arguments follow PETScVlasovSolver.pyx:27-42.  ``config`` is
    a ``Config`` (or anything with the same getters); ``da1``/``daph`` are
    accepted for signature parity and unused; ``grid`` is a ``Grid``.
    """

    def __init__(self, config, da1, daph, grid, H0, H1p, H1h, H2p, H2h, phi, charge=-1., coll_freq=0.,
                 coll_diff=1., coll_drag=1., regularisation=0., restart=30, context=None):
        if config.get_averaging_operator() != 'point':
            raise _lib.VVError("averaging operator %r is not provided" % config.get_averaging_operator())
        if regularisation != 0.:
            # Regularisation.function calls a method the reference does not define
            # (Regularisation.pyx:41-48): a non-zero epsilon fails there too
            raise _lib.VVError("regularisation is not provided")
        self.grid = grid
        self.charge = charge
        self.H0, self.H1p, self.H1h, self.H2p, self.H2h = H0, H1p, H1h, H2p, H2h
        self.phi = phi
        pc = effective_preconditioner(config.get_preconditioner())
        self.ctx = context if context is not None else Context(
            grid, bracket=config.get_poisson_bracket(), preconditioner=pc,
            laplace=config.get_laplace_operator(), charge=charge, restart=restart)
        self.poisson_bracket = PoissonBracket(config.get_poisson_bracket(), self.ctx)
        self.time_derivative = TimeDerivative(grid)
        # dissipation (:101-102): the context adds it to every residual and
        # Jacobian apply once set
        self.double_bracket = DoubleBracket.create(config.get_double_bracket(), None, grid, self.poisson_bracket,
                                                   coll_freq, context=self.ctx)
        self.collision_operator = CollisionOperator.create(config.get_collision_operator(), None, grid, coll_freq,
                                                           coll_diff, coll_drag, context=self.ctx)
        self.preconditioner = TensorProductPreconditioner(self.ctx, phi) if pc else None
        if self.preconditioner is not None:
            self.preconditioner.update_matrices()       # built from the initial phi (:104)
        self.Fp = self.ctx.new_field()
        self.Fh = self.ctx.new_field()
        self.Np = self.Up = self.Ep = self.Ap = None
        self.Nh = self.Uh = self.Eh = self.Ah = None

    def set_moments(self, Np, Up, Ep, Ap, Nh, Uh, Eh, Ah):
        """PETScVlasovSolver.pyx:123-132: the moment vectors the collision
        operator reads (their addresses; the caller refreshes the contents)."""
        self.Np, self.Up, self.Ep, self.Ap = Np, Up, Ep, Ap
        self.Nh, self.Uh, self.Eh, self.Ah = Nh, Uh, Eh, Ah
        self.ctx.set_moments(Np, Up, Ep, Ap, Nh, Uh, Eh, Ah)

    def update_history(self, F):
        """Fh = F; the preconditioner's matrices follow phi (:135-139)."""
        self.Fh.copy_(F)
        if self.preconditioner is not None:
            self.preconditioner.update_matrices()

    def update_previous(self, F):
        """Fp = F; H_ave from H0 + .5(H1p + H1h + H2p + H2h) (:142-157)."""
        self.Fp.copy_(F)
        self.ctx.set_hamiltonian(self.H0, self.H1p, self.H1h, self.H2p, self.H2h)

    # PETScNLVlasovMP.pyx:33-58 ----------------------------------------------
    def jacobian_solver(self, F, Y):
        self.ctx.jacobian_apply(F, Y)

    def function_solver(self, F, Y):
        self.ctx.residual(F, self.Fh, Y)

    # PETScVlasovSolver.pyx:160-198 ------------------------------------------
    def jacobian(self, F, Y):
        if self.preconditioner is None:
            self.ctx.jacobian_apply(F, Y)
        else:
            self.ctx.jacobian(F, Y)

    def function(self, F, Y):
        if self.preconditioner is None:
            self.ctx.residual(F, self.Fh, Y)
        else:
            self.ctx.function(F, self.Fh, Y)

    def snes_mult(self, snes, F, Y):
        self.jacobian(F, Y)

    def mult(self, mat, F, Y):
        self.jacobian(F, Y)

    def jacobian_mult(self, F, Y):
        self.jacobian(F, Y)

    def function_snes_mult(self, snes, F, Y):
        self.function(F, Y)

    def function_mult(self, F, Y):
        self.function(F, Y)

    def function_norm(self, F, Y):
        return self.ctx.function_norm(F, self.Fh, Y)

    def formJacobian(self, J):
        raise _lib.VVError("the assembled Jacobian (direct-solver path) is not provided")


class PETScPoissonSolver:
    """Periodic Poisson solver (PETScPoissonSolver.pyx:18-82 with the CFD4 /
    CFD2 stencils of PETScPoissonCFD4.pyx / PETScPoissonCFD2.pyx).

    ``solve(B, X)`` replaces the reference's CG + LU KSP on the singular
    matrix: an exact mean-free solve of A phi = b."""

    def __init__(self, dax, nx, hx, charge, laplace='cfd4', context=None, grid=None):
        self.nx, self.hx, self.charge = int(nx), float(hx), float(charge)
        self.hx2 = self.hx ** 2
        self.hx2_inv = 1. / self.hx2
        self.laplace = laplace
        if context is None:
            g = grid if grid is not None else Grid(nx, 8, nx * hx, 1.0)
            context = Context(g, preconditioner='', laplace=laplace, charge=charge)
        self.ctx = context

    def mult(self, mat, X, Y):
        self.matrix_mult(X, Y)

    def matrix_mult(self, X, Y):
        self.ctx.poisson_matrix_mult(X, Y)

    def function_mult(self, X, N, Y):
        self.ctx.poisson_function_mult(X, N, Y)

    def formRHS(self, N, B):
        """b = -(n - mean n) q  (PETScPoissonSolver.pyx:58-74)."""
        nmean = N.sum() / self.nx
        B.copy_(-(N - nmean) * self.charge)

    def solve_density(self, N, X):
        """KSPSolve of A x = formRHS(N): phi from the density."""
        self.ctx.poisson_solve(N, X)

    def formMat(self, A):
        raise _lib.VVError("the assembled Poisson matrix is not used: the solve is exact")


class Toolbox:
    """Moments and initialisation (Toolbox.pyx:38-224) on the GPU."""

    def __init__(self, da1, dax, grid, context):
        self.grid = grid
        self.ctx = context

    def compute_density(self, F, N):
        self.ctx.density(F, N)

    def compute_velocity_density(self, F, U):
        self.ctx.velocity_moments(F, U, self.ctx.new_x())

    def compute_energy_density(self, F, E):
        self.ctx.velocity_moments(F, self.ctx.new_x(), E)

    def compute_collision_factor(self, N, U, E, A):
        self.ctx.collision_factor(N, U, E, A)

    def potential_to_hamiltonian(self, P, H):
        self.ctx.potential_to_hamiltonian(P, H)

    def initialise_kinetic_hamiltonian(self, H, mass):
        H.copy_(torch.from_numpy(0.5 * self.grid.v ** 2 * mass))

    def initialise_distribution_function(self, F, init_function):
        from .core import initial_distribution
        F.copy_(torch.from_numpy(initial_distribution(self.grid, init_function)))


# ---------------------------------------------------------------------------
# PETSc KSP / SNES stand-ins for the objects the run script configures
# ---------------------------------------------------------------------------
class KSP:
    """GMRES on the MatPython context (the solver's ``mult``), PETSc defaults:
    restart = the context's restart (30), CGS, x0 = 0, dtol 1e5."""

    def __init__(self):
        self.rtol, self.atol, self.dtol, self.max_it = 1e-5, 1e-50, 1e5, 10000
        self.ctx = None
        self.type = 'gmres'
        self._res = None
        # 'double': the reference's FP64 GMRES.  'mixed': FP32 GMRES cycles,
        # refined either by the nonlinear (Picard) iteration around the solve
        # (refine='newton': each solve is one FP32 cycle to the relative
        # tolerance max(rtol, inner_rtol); the FP64 residual F, the update and
        # the Poisson solve of the outer loop drive the error to snes_atol) or
        # inside the solve (refine='ksp': FP64 iterative refinement on the true
        # preconditioned residual to the KSP's own rtol).
        # fp64_above (refine='newton'): a Picard solve whose starting residual
        # exceeds it runs in FP64 at the same relative tolerance.  The FP32
        # cycle leaves broadband rounding noise of about 2^-24 times the
        # correction in f, which GMRES removes slowly in every later solve of
        # the step; where that noise would come within a factor 16 of the
        # nonlinear tolerance (residual > snes_atol * 2^20, the run script's
        # choice) the solve is FP64.  None: every solve FP32.
        self.precision = 'double'
        self.refine = 'newton'
        self.fp64_above = None
        self.inner_rtol = 1e-5
        self.max_cycles = 50
        self.cycles = 0
        self.last_precision = 'double'

    def setPrecision(self, precision, inner_rtol=None, max_cycles=None, refine=None, fp64_above=False):
        """fp64_above: residual above which a mixed solve runs in FP64 (None:
        never; the default False leaves it unchanged)."""
        if precision not in ('double', 'mixed'):
            raise _lib.VVError("precision must be 'double' or 'mixed'")
        if refine is not None and refine not in ('newton', 'ksp'):
            raise _lib.VVError("refine must be 'newton' or 'ksp'")
        self.precision = precision
        if refine is not None:
            self.refine = refine
        if fp64_above is not False:
            self.fp64_above = None if fp64_above is None else float(fp64_above)
        if inner_rtol is not None:
            self.inner_rtol = float(inner_rtol)
        if max_cycles is not None:
            self.max_cycles = int(max_cycles)

    def fp64_solve(self, residual):
        """Whether a mixed-precision solve starting from this Picard residual
        runs in FP64 (refine='newton' only; None: residual unknown)."""
        return (self.refine == 'newton' and self.fp64_above is not None and residual is not None
                and residual > self.fp64_above)

    def mixed_args(self):
        """(rtol, atol, dtol, max_it, inner_rtol, max_cycles, fp64_check) of a mixed solve."""
        if self.refine == 'newton':
            return (max(self.rtol, self.inner_rtol), self.atol, self.dtol, self.max_it, self.inner_rtol,
                    self.max_cycles, False)
        return (self.rtol, self.atol, self.dtol, self.max_it, self.inner_rtol, self.max_cycles, True)

    def setType(self, t):
        if t != 'gmres':
            raise _lib.VVError("only KSP type gmres is provided")
        self.type = t

    def setTolerances(self, rtol=None, atol=None, divtol=None, max_it=None):
        if rtol is not None:
            self.rtol = rtol
        if atol is not None:
            self.atol = atol
        if divtol is not None:
            self.dtol = divtol
        if max_it is not None:
            self.max_it = max_it

    def setOperators(self, A, P=None):
        ctx = getattr(A, 'ctx', None)
        if ctx is None:
            raise _lib.VVError("KSP operator must be a PETScVlasovSolver (its mult runs on the GPU)")
        self.ctx = ctx

    def solve(self, b, x):
        if self.precision == 'mixed':
            self._res, self.cycles = self.ctx.ksp_solve_mixed(b, x, *self.mixed_args())
        else:
            self._res = self.ctx.ksp_solve(b, x, self.rtol, self.atol, self.dtol, self.max_it)

    def getIterationNumber(self):
        return self._res.its if self._res else 0

    def getConvergedReason(self):
        return self._res.reason if self._res else 0

    def getResidualNorm(self):
        return self._res.rnorm if self._res else 0.0


class SNES:
    """SNES of type 'ksponly' with the solver's function and matrix-free
    Jacobian (run_nonlinear_matrixfree_split_mp.py:47-53)."""

    def __init__(self):
        self.ksp = KSP()
        self.solver = None
        self.type = 'ksponly'
        self.linear_its = 0
        self.last = None
        # Picard residual the next solve starts from (set by the run script;
        # None: unknown), for the mixed-precision KSP's choice of precision
        self.picard_residual = None

    def setType(self, t):
        if t != 'ksponly':
            raise _lib.VVError("only SNES type ksponly is provided")

    def setFunction(self, function, fb):
        solver = getattr(function, '__self__', None)
        if not isinstance(solver, PETScVlasovSolver):
            raise _lib.VVError("the SNES function must be PETScVlasovSolver.function_snes_mult")
        self.solver = solver
        self.ksp.setOperators(solver)

    def setJacobian(self, update, J, P=None):
        pass

    def setFromOptions(self):
        pass

    def getKSP(self):
        return self.ksp

    def solve(self, b, x):
        """SNESSolve_KSPONLY: F = P^-1 F(x) (F(x) without a preconditioner);
        KSPSolve; x -= delta unless the linear solve diverged."""
        s = self.solver
        k = self.ksp
        if k.precision == 'mixed' and k.fp64_solve(self.picard_residual):
            self.last = s.ctx.snes_ksponly(x, s.Fh, max(k.rtol, k.inner_rtol), k.atol, k.dtol, k.max_it)
            k.cycles = 0
            k.last_precision = 'double'
        elif k.precision == 'mixed':
            self.last, k.cycles = s.ctx.snes_ksponly_mixed(x, s.Fh, *k.mixed_args())
            k.last_precision = 'mixed'
        else:
            self.last = s.ctx.snes_ksponly(x, s.Fh, k.rtol, k.atol, k.dtol, k.max_it)
            k.last_precision = 'double'
        k._res = self.last
        self.linear_its = self.last.its

    def getLinearSolveIterations(self):
        return self.linear_its
