"""The C-ABI library loads on the CPU and exports every function include/vv1d.h
declares; argument validation runs without a GPU (no compute calls here)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vv1d.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vv_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from vivlasov1d_amd import _lib
    if not os.path.exists(_lib.LIBPATH):
        from vivlasov1d_amd import build
        build.build()
    return _lib.load()


def test_header_declares_the_api():
    names = declared_functions()
    for n in ("vv_ctx_create", "vv_residual", "vv_jacobian_apply", "vv_precond_apply", "vv_ksp_solve",
              "vv_snes_ksponly", "vv_moments", "vv_poisson_solve", "vv_density"):
        assert n in names


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    from vivlasov1d_amd import _lib
    assert set(declared_functions()) == set(_lib.SIGNATURES)


@pytest.mark.parametrize("field,value,msg", [
    ("nx", 5000, "nx must be"), ("nv", 4, "nv must be"), ("bracket", 3, "bracket"),
    ("laplace", 6, "laplace"), ("restart", 0, "restart"), ("precond", 7, "preconditioner")])
def test_create_rejects_bad_parameters(lib, field, value, msg):
    from vivlasov1d_amd import _lib
    import numpy as np
    v = np.linspace(-1, 1, 64)
    eig = np.zeros(2 * 64)
    p = _lib.VVParams(64, 64, 0.1, 0.1, 0.05, -1.0, 4, 1, 4, 30,
                      v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                      eig.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    setattr(p, field, value)
    out = ctypes.c_void_p()
    ret = lib.vv_ctx_create(ctypes.byref(p), None, ctypes.byref(out))
    assert ret == -1 or ret == -3
    assert out.value is None
    assert msg in lib.vv_last_error(None).decode()


def test_null_context_is_an_error(lib):
    assert lib.vv_residual(None, None, None, None) == -1
    assert lib.vv_synchronize(None) == -1
    assert lib.vv_batched_timestep(None, None, None, 1e-10, 20, None, None, 0, None) == -1


def test_batched_timestep_layouts():
    """vv_batch_fields: twelve pointers; vv_instance_step: int32 picard,
    int32 reverted, int64 gmres, double initial, double final."""
    from vivlasov1d_amd import _lib
    F, S = _lib.BatchFields, _lib.InstanceStep
    assert ctypes.sizeof(F) == 12 * ctypes.sizeof(ctypes.c_void_p)
    assert [f[0] for f in F._fields_] == ["f", "fh", "fl", "fy", "h0", "h1", "h1h", "h2", "h2h", "n", "phi", "phih"]
    assert ctypes.sizeof(S) == 32
    assert (S.picard.offset, S.reverted.offset, S.gmres.offset, S.initial.offset, S.final.offset) == (0, 4, 8, 16, 24)


def test_mixed_params_layout():
    """vv_mixed_params: double inner_rtol; int32 max_cycles; int32 fp64_check."""
    from vivlasov1d_amd import _lib
    M = _lib.MixedParams
    assert ctypes.sizeof(M) == 16
    assert (M.inner_rtol.offset, M.max_cycles.offset, M.fp64_check.offset) == (0, 8, 12)


def test_ksp_mixed_refinement_arguments():
    """refine='newton' solves each system by one FP32 cycle to max(rtol,
    inner_rtol) with no FP64 check (the Picard loop refines); refine='ksp'
    keeps the KSP rtol and refines on the FP64 residual."""
    from vivlasov1d_amd import _lib
    from vivlasov1d_amd.solvers import KSP
    k = KSP()
    k.setTolerances(rtol=1e-10, atol=1e-50, max_it=1000)
    k.setPrecision('mixed', inner_rtol=1e-5)
    assert k.mixed_args() == (1e-5, 1e-50, 1e5, 1000, 1e-5, 50, False)
    k.setTolerances(rtol=1e-3)
    assert k.mixed_args()[0] == 1e-3
    k.setTolerances(rtol=1e-10)
    k.setPrecision('mixed', refine='ksp', max_cycles=7)
    assert k.mixed_args() == (1e-10, 1e-50, 1e5, 1000, 1e-5, 7, True)
    with pytest.raises(_lib.VVError):
        k.setPrecision('mixed', refine='sometimes')
    with pytest.raises(_lib.VVError):
        k.setPrecision('half')


def test_dissipation_layout_and_kinds():
    """vv_dissipation: int32 kind, int32 pad, double coll_freq, coll_diff,
    coll_drag; the kind codes match include/vv1d.h and the reference's
    create() names (CollisionOperator.pyx:52-57, DoubleBracket.pyx:51-56)."""
    from vivlasov1d_amd import _lib
    D = _lib.Dissipation
    assert ctypes.sizeof(D) == 32
    assert (D.kind.offset, D.coll_freq.offset, D.coll_diff.offset, D.coll_drag.offset) == (0, 8, 16, 24)
    text = open(HEADER).read()
    for name, code in (("COLLT", 1), ("COLLE", 2), ("FFH", 3), ("FHH", 4)):
        assert "#define VV_DISSIPATION_%s %d" % (name, code) in " ".join(text.split())
        assert _lib.DISSIPATION[name.lower()] == code


def test_config_selects_dissipation():
    """config.py:148-163: collisions only with dissipation_type 'collisions'
    and coll_freq != 0; double brackets only with 'double_bracket'."""
    from vivlasov1d_amd.core import Config
    cfg = Config({'solver': {'dissipation_type': 'collisions', 'collision_operator': 'collT', 'coll_freq': 1e-3}})
    assert cfg.get_collision_operator() == 'collt' and cfg.get_double_bracket() == ''
    cfg = Config({'solver': {'dissipation_type': 'double_bracket', 'bracket_operator': 'FHH', 'coll_freq': 1e-3}})
    assert cfg.get_double_bracket() == 'fhh' and cfg.get_collision_operator() == ''
    cfg = Config({'solver': {'dissipation_type': 'collisions', 'collision_operator': 'collE', 'coll_freq': 0.0}})
    assert cfg.get_collision_operator() == '' and cfg.is_dissipation_none()


def test_host_classes_keep_the_reference_method_names():
    """The Python layer's classes carry the reference's callbacks
    (PETScVlasovSolver.pyx:130-198, TensorProduct.pyx:66-160,
    PETScPoissonSolver.pyx:58-100)."""
    from vivlasov1d_amd import solvers
    need = {
        solvers.PETScVlasovSolver: ['set_moments', 'update_history', 'update_previous', 'snes_mult', 'mult',
                                    'jacobian_mult', 'function_snes_mult', 'function_mult', 'function_norm',
                                    'jacobian', 'function'],
        solvers.TensorProductPreconditioner: ['create', 'tensorProduct', 'update_matrices'],
        solvers.PETScPoissonSolver: ['formRHS', 'mult', 'function_mult'],
    }
    for cls, names in need.items():
        for n in names:
            assert callable(getattr(cls, n, None)), (cls.__name__, n)
