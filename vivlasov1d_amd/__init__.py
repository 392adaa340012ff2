"""MI355X (gfx950) implicit timestep of the 1D1V Vlasov-Poisson variational
integrator: Arakawa J1/J2/J4 residual and Jacobian, kinetic FFT(x) x
tridiagonal(v) preconditioner, GMRES and the periodic Poisson solve, as
hand-written HIP kernels behind the C ABI of libvv1d.so.

The Python classes keep the interface of the reference's solver objects
(``PETScVlasovSolver``, ``PETScPoissonSolver``, ``Toolbox``, the SNES/KSP
callbacks) and its run script (``Runscript``).
"""

from .core import (Config, Grid, bump_on_tail, initial_distribution, kinetic_eigenvalues,
                   kinetic_hamiltonian, landau, maxwellian, normalise, twostream)

__all__ = ["Config", "Grid", "bump_on_tail", "initial_distribution", "kinetic_eigenvalues",
           "kinetic_hamiltonian", "landau", "maxwellian", "normalise", "twostream"]


def __getattr__(name):
    # the GPU-backed classes import torch and the HIP library lazily
    if name in ("Context", "PETScVlasovSolver", "PETScPoissonSolver", "TensorProductPreconditioner",
                "PoissonBracket", "TimeDerivative", "Toolbox", "KSP", "SNES"):
        from . import solvers
        return getattr(solvers, name)
    if name in ("Runscript", "ResidualTooLarge"):
        from . import runner
        return getattr(runner, name)
    raise AttributeError(name)
