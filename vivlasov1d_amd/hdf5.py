"""Run output in the reference's HDF5 schema, written through the HDF5 C
library with ctypes (h5py is not part of this stack).

The reference (run_base.py:528-562, 766-783) writes, through PETSc's HDF5
viewer, one file per run with

* attribute ``charge`` on the root group;
* ``x`` (nx,) and ``v`` (nv,): the grid coordinates, written once;
* per saved timestep, with a leading time index: ``t`` (T, 1, 1);
  ``h0``, ``h1``, ``h2``, ``f`` (T, nv, nx) -- 2D fields stored [t][v][x];
  ``phi_int``, ``phi_ext``, ``n``, ``u``, ``e`` (T, nx).

Its readers rely on exactly these shapes: ``hdf5['t'][:,0,0]``,
``hdf5['x'][:]`` (Grid.pyx:53-59), ``hdf5['f'][iTime,:,:].T``
(distribution_function.pyx:319-324), ``hdf5['h0'][iTime,:,:].T``
(hamiltonian.pyx:159-166), ``hdf5['phi_int'][iTime,:]`` (potential.pyx:64-66).
Time-indexed datasets are extendible along the first axis, chunked one
timestep per chunk, FP64 little-endian.

The library is found through the ``VV1D_HDF5_LIB`` environment variable, the
linker search path, or an HDF5 shipped with the image's conda tree; without
one, creating an :class:`HDF5Output` raises (runs without output do not need
it).
"""

import ctypes
import ctypes.util
import glob
import os

import numpy as np

_hid = ctypes.c_int64
_herr = ctypes.c_int
_hsize = ctypes.c_uint64
_HS = ctypes.POINTER(_hsize)
H5F_ACC_RDONLY = 0x0000
H5F_ACC_TRUNC = 0x0002
H5F_ACC_RDWR = 0x0001
H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SCALAR = 0
H5S_SELECT_SET = 0
H5S_UNLIMITED = (1 << 64) - 1

TIME_SERIES_2D = ("h0", "h1", "h2", "f")
TIME_SERIES_1D = ("phi_int", "phi_ext", "n", "u", "e")

_LIB = None


class HDF5Error(RuntimeError):
    pass


def _candidates():
    env = os.environ.get("VV1D_HDF5_LIB")
    if env:
        yield env
    name = ctypes.util.find_library("hdf5")
    if name:
        yield name
    for pat in ("/opt/conda/lib/libhdf5.so", "/opt/conda/lib/libhdf5.so.*", "/usr/lib/x86_64-linux-gnu/libhdf5*.so*"):
        for p in sorted(glob.glob(pat)):
            if "_hl" not in p and "fortran" not in p and "cpp" not in p:
                yield p


def library():
    """The loaded HDF5 C library (raises HDF5Error when none is found)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    errors = []
    for cand in _candidates():
        try:
            lib = ctypes.CDLL(cand)
            break
        except OSError as e:
            errors.append("%s: %s" % (cand, e))
    else:
        raise HDF5Error("no HDF5 C library found (set VV1D_HDF5_LIB); tried: %s" % "; ".join(errors))
    sig = {
        "H5open": (_herr, []),
        "H5Fcreate": (_hid, [ctypes.c_char_p, ctypes.c_uint, _hid, _hid]),
        "H5Fopen": (_hid, [ctypes.c_char_p, ctypes.c_uint, _hid]),
        "H5Fclose": (_herr, [_hid]),
        "H5Fflush": (_herr, [_hid, ctypes.c_int]),
        "H5Screate": (_hid, [ctypes.c_int]),
        "H5Screate_simple": (_hid, [ctypes.c_int, _HS, _HS]),
        "H5Sclose": (_herr, [_hid]),
        "H5Sselect_hyperslab": (_herr, [_hid, ctypes.c_int, _HS, _HS, _HS, _HS]),
        "H5Sget_simple_extent_ndims": (ctypes.c_int, [_hid]),
        "H5Sget_simple_extent_dims": (ctypes.c_int, [_hid, _HS, _HS]),
        "H5Pcreate": (_hid, [_hid]),
        "H5Pset_chunk": (_herr, [_hid, ctypes.c_int, _HS]),
        "H5Pclose": (_herr, [_hid]),
        "H5Dcreate2": (_hid, [_hid, ctypes.c_char_p, _hid, _hid, _hid, _hid, _hid]),
        "H5Dopen2": (_hid, [_hid, ctypes.c_char_p, _hid]),
        "H5Dset_extent": (_herr, [_hid, _HS]),
        "H5Dget_space": (_hid, [_hid]),
        "H5Dwrite": (_herr, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
        "H5Dread": (_herr, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
        "H5Dclose": (_herr, [_hid]),
        "H5Acreate2": (_hid, [_hid, ctypes.c_char_p, _hid, _hid, _hid, _hid]),
        "H5Aopen": (_hid, [_hid, ctypes.c_char_p, _hid]),
        "H5Awrite": (_herr, [_hid, _hid, ctypes.c_void_p]),
        "H5Aread": (_herr, [_hid, _hid, ctypes.c_void_p]),
        "H5Aclose": (_herr, [_hid]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.H5open() < 0:
        raise HDF5Error("H5open failed")
    _LIB = lib
    return lib


def _global(lib, name):
    return _hid.in_dll(lib, name).value


def _dims(vals):
    return (_hsize * len(vals))(*vals)


def _ok(ret, what):
    if ret < 0:
        raise HDF5Error("HDF5 call failed: %s" % what)
    return ret


class HDF5Output:
    """One run's output file in the reference's layout.

    ``save(t, h0, h1, h2, f, phi_int, phi_ext, n, u, e)`` appends one
    timestep; 2D fields are (nv, nx) arrays in natural order (x contiguous),
    x-vectors (nx,).  Arrays may be CUDA tensors (copied to the host here).
    """

    def __init__(self, path, x, v, charge):
        lib = self.lib = library()
        self.f64 = _global(lib, "H5T_NATIVE_DOUBLE_g")
        self.f64le = _global(lib, "H5T_IEEE_F64LE_g")
        self.dcpl_cls = _global(lib, "H5P_CLS_DATASET_CREATE_ID_g")
        self.nx, self.nv = len(x), len(v)
        self.fid = _ok(lib.H5Fcreate(os.fsencode(path), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "create " + path)
        self.path = path
        self.steps = 0
        self.dsets = {}
        # attribute charge (h5py attrs on the root group, run_base.py:534-536)
        sp = _ok(lib.H5Screate(H5S_SCALAR), "scalar space")
        at = _ok(lib.H5Acreate2(self.fid, b"charge", self.f64le, sp, H5P_DEFAULT, H5P_DEFAULT), "attribute")
        c = ctypes.c_double(float(charge))
        _ok(lib.H5Awrite(at, self.f64, ctypes.byref(c)), "write charge")
        lib.H5Aclose(at)
        lib.H5Sclose(sp)
        self._write_static("x", np.asarray(x, dtype=np.float64))
        self._write_static("v", np.asarray(v, dtype=np.float64))
        self._create_series("t", (1, 1))
        for name in TIME_SERIES_2D:
            self._create_series(name, (self.nv, self.nx))
        for name in TIME_SERIES_1D:
            self._create_series(name, (self.nx,))

    def _write_static(self, name, arr):
        lib = self.lib
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        sp = _ok(lib.H5Screate_simple(arr.ndim, _dims(arr.shape), None), "space")
        ds = _ok(lib.H5Dcreate2(self.fid, name.encode(), self.f64le, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                 "create " + name)
        _ok(lib.H5Dwrite(ds, self.f64, H5S_ALL, H5S_ALL, H5P_DEFAULT, arr.ctypes.data), "write " + name)
        lib.H5Dclose(ds)
        lib.H5Sclose(sp)

    def _create_series(self, name, shape):
        lib = self.lib
        dims = (0,) + tuple(shape)
        maxd = (H5S_UNLIMITED,) + tuple(shape)
        sp = _ok(lib.H5Screate_simple(len(dims), _dims(dims), _dims(maxd)), "space")
        pl = _ok(lib.H5Pcreate(self.dcpl_cls), "dataset plist")
        _ok(lib.H5Pset_chunk(pl, len(dims), _dims((1,) + tuple(shape))), "chunk")
        ds = _ok(lib.H5Dcreate2(self.fid, name.encode(), self.f64le, sp, H5P_DEFAULT, pl, H5P_DEFAULT),
                 "create " + name)
        lib.H5Pclose(pl)
        lib.H5Sclose(sp)
        self.dsets[name] = (ds, tuple(shape))

    def _append(self, name, arr, k):
        lib = self.lib
        ds, shape = self.dsets[name]
        if hasattr(arr, "detach"):
            arr = arr.detach().cpu().numpy()
        arr = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).reshape(shape))
        _ok(lib.H5Dset_extent(ds, _dims((k + 1,) + shape)), "extend " + name)
        fsp = _ok(lib.H5Dget_space(ds), "space")
        _ok(lib.H5Sselect_hyperslab(fsp, H5S_SELECT_SET, _dims((k,) + (0,) * len(shape)), None,
                                    _dims((1,) + shape), None), "select")
        msp = _ok(lib.H5Screate_simple(len(shape) + 1, _dims((1,) + shape), None), "memory space")
        _ok(lib.H5Dwrite(ds, self.f64, msp, fsp, H5P_DEFAULT, arr.ctypes.data), "write " + name)
        lib.H5Sclose(msp)
        lib.H5Sclose(fsp)

    def save(self, t, h0, h1, h2, f, phi_int, phi_ext, n, u, e):
        k = self.steps
        self._append("t", np.array([t], dtype=np.float64), k)
        for name, arr in zip(TIME_SERIES_2D, (h0, h1, h2, f)):
            self._append(name, arr, k)
        for name, arr in zip(TIME_SERIES_1D, (phi_int, phi_ext, n, u, e)):
            self._append(name, arr, k)
        self.steps += 1
        self.lib.H5Fflush(self.fid, 0)

    def close(self):
        if getattr(self, "fid", None) is None:
            return
        for ds, _ in self.dsets.values():
            self.lib.H5Dclose(ds)
        self.lib.H5Fclose(self.fid)
        self.fid = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read(path):
    """All datasets of a file written by HDF5Output, as {name: ndarray}, plus
    the root attributes under ``'attrs'``."""
    lib = library()
    f64 = _global(lib, "H5T_NATIVE_DOUBLE_g")
    fid = _ok(lib.H5Fopen(os.fsencode(path), H5F_ACC_RDONLY, H5P_DEFAULT), "open " + path)
    out = {}
    try:
        for name in ("x", "v", "t") + TIME_SERIES_2D + TIME_SERIES_1D:
            ds = lib.H5Dopen2(fid, name.encode(), H5P_DEFAULT)
            if ds < 0:
                continue
            sp = lib.H5Dget_space(ds)
            nd = lib.H5Sget_simple_extent_ndims(sp)
            dims = (_hsize * nd)()
            lib.H5Sget_simple_extent_dims(sp, dims, None)
            arr = np.empty(tuple(int(d) for d in dims), dtype=np.float64)
            _ok(lib.H5Dread(ds, f64, H5S_ALL, H5S_ALL, H5P_DEFAULT, arr.ctypes.data), "read " + name)
            out[name] = arr
            lib.H5Sclose(sp)
            lib.H5Dclose(ds)
        at = lib.H5Aopen(fid, b"charge", H5P_DEFAULT)
        if at >= 0:
            c = ctypes.c_double()
            lib.H5Aread(at, f64, ctypes.byref(c))
            lib.H5Aclose(at)
            out["attrs"] = {"charge": c.value}
    finally:
        lib.H5Fclose(fid)
    return out
