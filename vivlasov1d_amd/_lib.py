"""ctypes binding of libvv1d.so (the C ABI declared in include/vv1d.h).

The library is loaded from this package directory.  There is no fallback:
if it has not been built, or the GPU runtime cannot be loaded, every
operation raises.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libvv1d.so")

VV_OK = 0
ERRORS = {-1: "invalid argument", -2: "HIP runtime error", -3: "unsupported option", -4: "non-finite value"}

BRACKETS = {"arakawaj1": 1, "arakawaj2": 2, "arakawaj4": 4}
# the _scipy schemes of the reference state the same operators with SciPy's
# banded solver (TensorProduct.pyx:66-86)
PRECONDITIONERS = {"": 0, None: 0, "none": 0, "tensor_kinetic": 1, "tensor_kinetic_scipy": 1,
                   "tensor_potential": 2, "tensor_potential_scipy": 2}
LAPLACE = {"cfd2": 2, "cfd4": 4}
LINEAR = {"A": 0, "A0": 1, "AT": 2, "A0T": 3}
# dissipation terms: collision operators (collision_operator collT / collE) and
# double brackets (bracket_operator ffh / fhh), CollisionOperator.pyx:43-57,
# DoubleBracket.pyx:43-56
DISSIPATION = {"": 0, None: 0, "none": 0, "collt": 1, "colle": 2, "ffh": 3, "fhh": 4}
RK4_RULES = {"rk4_16": 1, "rk4_18": 2}

KSP_REASONS = {2: "CONVERGED_RTOL", 3: "CONVERGED_ATOL", -3: "DIVERGED_ITS", -4: "DIVERGED_DTOL",
               -5: "DIVERGED_BREAKDOWN", -9: "DIVERGED_NANORINF", 0: "ITERATING"}


class VVParams(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("nv", ctypes.c_int32),
                ("hx", ctypes.c_double), ("hv", ctypes.c_double), ("ht", ctypes.c_double),
                ("charge", ctypes.c_double),
                ("bracket", ctypes.c_int32), ("precond", ctypes.c_int32), ("laplace", ctypes.c_int32),
                ("restart", ctypes.c_int32),
                ("v", ctypes.POINTER(ctypes.c_double)), ("eigen", ctypes.POINTER(ctypes.c_double))]


class KspParams(ctypes.Structure):
    _fields_ = [("rtol", ctypes.c_double), ("atol", ctypes.c_double), ("dtol", ctypes.c_double),
                ("max_it", ctypes.c_int32)]


class KspResult(ctypes.Structure):
    _fields_ = [("its", ctypes.c_int32), ("reason", ctypes.c_int32), ("rnorm", ctypes.c_double),
                ("rnorm0", ctypes.c_double)]


class BatchFields(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("f", "fh", "fl", "fy", "h0", "h1", "h1h", "h2", "h2h", "n", "phi",
                                               "phih")]


class InstanceStep(ctypes.Structure):
    _fields_ = [("picard", ctypes.c_int32), ("reverted", ctypes.c_int32), ("gmres", ctypes.c_int64),
                ("initial", ctypes.c_double), ("final", ctypes.c_double)]


class MixedParams(ctypes.Structure):
    _fields_ = [("inner_rtol", ctypes.c_double), ("max_cycles", ctypes.c_int32), ("fp64_check", ctypes.c_int32)]


class Dissipation(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("pad", ctypes.c_int32), ("coll_freq", ctypes.c_double),
                ("coll_diff", ctypes.c_double), ("coll_drag", ctypes.c_double)]


_P = ctypes.c_void_p
_D = ctypes.c_double
_I64 = ctypes.c_int64
_PD = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes)
SIGNATURES = {
    "vv_ctx_create": (ctypes.c_int, [ctypes.POINTER(VVParams), _P, ctypes.POINTER(_P)]),
    "vv_ctx_create_batched": (ctypes.c_int, [ctypes.POINTER(VVParams), ctypes.c_int32, _PD, _P,
                                             ctypes.POINTER(_P)]),
    "vv_batch_size": (ctypes.c_int32, [_P]),
    "vv_set_active": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32)]),
    "vv_set_graph": (ctypes.c_int, [_P, ctypes.c_int]),
    "vv_set_tuning": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32]),
    "vv_get_tuning": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    "vv_autotune": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    "vv_set_fusion": (ctypes.c_int, [_P, ctypes.c_int]),
    "vv_ctx_destroy": (ctypes.c_int, [_P]),
    "vv_last_error": (ctypes.c_char_p, [_P]),
    "vv_set_stream": (ctypes.c_int, [_P, _P]),
    "vv_synchronize": (ctypes.c_int, [_P]),
    "vv_scratch_bytes": (ctypes.c_int64, [_P]),
    "vv_set_hamiltonian": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "vv_residual": (ctypes.c_int, [_P, _P, _P, _P]),
    "vv_jacobian_apply": (ctypes.c_int, [_P, _P, _P]),
    "vv_bracket_add": (ctypes.c_int, [_P, _P, _P, _D]),
    "vv_precond_apply": (ctypes.c_int, [_P, _P, _P]),
    "vv_precond_update": (ctypes.c_int, [_P, _P]),
    "vv_function": (ctypes.c_int, [_P, _P, _P, _P]),
    "vv_jacobian": (ctypes.c_int, [_P, _P, _P]),
    "vv_function_norm": (ctypes.c_int, [_P, _P, _P, _P, _PD]),
    "vv_residual_norms": (ctypes.c_int, [_P, _P, _P, _P, _PD, _PD]),
    "vv_density": (ctypes.c_int, [_P, _P, _P]),
    "vv_velocity_moments": (ctypes.c_int, [_P, _P, _P, _P]),
    "vv_set_dissipation": (ctypes.c_int, [_P, ctypes.POINTER(Dissipation)]),
    "vv_set_moments": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "vv_collision_factor": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "vv_collision_add": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _D]),
    "vv_double_bracket_function": (ctypes.c_int, [_P, _P, _P, _D]),
    "vv_double_bracket_jacobian": (ctypes.c_int, [_P, _P, _P, _P, _D]),
    "vv_symplectic_kinetic": (ctypes.c_int, [_P, _P, _P, _D, ctypes.c_int32]),
    "vv_symplectic_potential": (ctypes.c_int, [_P, _P, _P, _P, _D, ctypes.c_int32]),
    "vv_rk4": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.c_int32, ctypes.c_int32]),
    "vv_poisson_solve": (ctypes.c_int, [_P, _P, _P]),
    "vv_poisson_matrix_mult": (ctypes.c_int, [_P, _P, _P]),
    "vv_poisson_function_mult": (ctypes.c_int, [_P, _P, _P, _P]),
    "vv_potential_to_hamiltonian": (ctypes.c_int, [_P, _P, _P]),
    "vv_moments": (ctypes.c_int, [_P, _P, _P, _P, _P, _PD]),
    "vv_moments_noalloc": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "vv_norm2": (ctypes.c_int, [_P, _P, _I64, _PD]),
    "vv_dot": (ctypes.c_int, [_P, _P, _P, _I64, _PD]),
    "vv_axpy": (ctypes.c_int, [_P, _D, _P, _P, _I64]),
    "vv_ksp_solve": (ctypes.c_int, [_P, _P, _P, ctypes.POINTER(KspParams), ctypes.POINTER(KspResult)]),
    "vv_snes_ksponly": (ctypes.c_int, [_P, _P, _P, ctypes.POINTER(KspParams), ctypes.POINTER(KspResult)]),
    "vv_batched_timestep": (ctypes.c_int, [_P, ctypes.POINTER(BatchFields), ctypes.POINTER(KspParams), ctypes.c_double,
                                           ctypes.c_int32, ctypes.POINTER(InstanceStep), ctypes.POINTER(ctypes.c_double),
                                           ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    "vv_ksp_solve_mixed": (ctypes.c_int, [_P, _P, _P, ctypes.POINTER(KspParams), ctypes.POINTER(MixedParams),
                                          ctypes.POINTER(KspResult), ctypes.POINTER(ctypes.c_int32)]),
    "vv_snes_ksponly_mixed": (ctypes.c_int, [_P, _P, _P, ctypes.POINTER(KspParams), ctypes.POINTER(MixedParams),
                                             ctypes.POINTER(KspResult), ctypes.POINTER(ctypes.c_int32)]),
    "vv_jacobian_apply32": (ctypes.c_int, [_P, _P, _P]),
    "vv_precond_apply32": (ctypes.c_int, [_P, _P, _P]),
    "vv_linear_apply": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_int32]),
    "vv_precond_apply_t": (ctypes.c_int, [_P, _P, _P]),
    "vv_adjoint_solve": (ctypes.c_int, [_P, _P, _P, _P, ctypes.POINTER(KspParams), ctypes.POINTER(KspResult)]),
    "vv_adjoint_step": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.POINTER(KspParams), ctypes.POINTER(KspResult)]),
    "vv_invariants": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _PD]),
    "vv_profile_enable": (ctypes.c_int, [_P, ctypes.c_int]),
    "vv_profile_sample": (ctypes.c_int, [_P, ctypes.c_int]),
    "vv_profile_read": (ctypes.c_int, [_P, ctypes.c_char_p, _PD, ctypes.POINTER(ctypes.c_int64), _PD]),
}

_lib = None


class VVError(RuntimeError):
    pass


def load():
    """Load libvv1d.so once; raise if it is missing (there is no CPU path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise VVError("libvv1d.so is not built (%s); run `python -m vivlasov1d_amd.build`" % LIBPATH)
    lib = ctypes.CDLL(LIBPATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(ret, ctx=None):
    if ret != VV_OK:
        lib = load()
        msg = lib.vv_last_error(ctx)
        msg = msg.decode() if msg else ""
        raise VVError("libvv1d: %s (%d): %s" % (ERRORS.get(ret, "error"), ret, msg))
    return ret
