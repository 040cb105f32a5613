"""Python binding of libsatt_h5.so (include/satt_h5.h): the reference's HDF5
model and kernel files (src/acqui_2d.f90:210-264, src/tomo.f90:72-98,350-385,
src/model.f90:95-105).

Arrays are numpy arrays in the reference's Fortran shapes (order "F"): the
2-D model svel (nperiod, nx, ny), the 3-D model (nx, ny, nz), the gradient
(nker, nx, ny, nz).  The files hold transpose_k of them, so a reader in C
order (h5py, h5dump) sees /vel as (nperiod, nx, ny) with vel[p, i, j] =
svel[p, i, j].
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsatt_h5.so")
OK, E_ARG, E_H5 = 0, -1, -9
MAX_RANK = 4

_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


class H5Error(RuntimeError):
    def __init__(self, status, what):
        super().__init__(f"{what}: status {status}")
        self.status = status


def load(path=LIB_PATH):
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise H5Error(E_ARG, f"{path} not built (the image's HDF5 library is needed)")
        _lib = ctypes.CDLL(path)
        for n in ("satt_h5_open", "satt_h5_close", "satt_h5_exist", "satt_h5_add",
                  "satt_h5_add_transposed", "satt_h5_shape", "satt_h5_get",
                  "satt_h5_write_model_2d", "satt_h5_write_iter_2d",
                  "satt_h5_write_model_3d_init", "satt_h5_write_model_3d_iter",
                  "satt_h5_write_gradient_3d", "satt_h5_write_kdensity_3d",
                  "satt_h5_read_model_3d"):
            getattr(_lib, n).restype = ctypes.c_int
    return _lib


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _f(a):
    return None if a is None else np.asfortranarray(a, dtype=np.float64)


def _c(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _chk(st, what):
    if st < 0:
        raise H5Error(st, what)
    return st


def _dims(shape):
    return (ctypes.c_longlong * MAX_RANK)(*[int(v) for v in shape])


class File:
    """hdf5_file (src/shared/hdf5_interface.f90): open, add, get, close."""

    def __init__(self, path, status="old", action="rw"):
        self.lib = load()
        p = ctypes.c_void_p()
        _chk(self.lib.satt_h5_open(ctypes.byref(p), path.encode(), status.encode(),
                                   action.encode()), f"open {path}")
        self.ptr = p

    def close(self):
        if self.ptr:
            _chk(self.lib.satt_h5_close(self.ptr), "close")
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def exist(self, name):
        return self.lib.satt_h5_exist(self.ptr, name.encode()) == 1

    def add(self, name, value, transposed=False):
        """hf%add(name, value), or hf%add(name, transpose_k(value))."""
        v = _f(value)
        v = v.reshape(v.shape or (1,), order="F")
        fn = self.lib.satt_h5_add_transposed if transposed else self.lib.satt_h5_add
        _chk(fn(self.ptr, name.encode(), v.ndim, _dims(v.shape), _d(v)), f"add {name}")

    def shape(self, name):
        """The dataset's Fortran shape."""
        r = ctypes.c_int(0)
        d = (ctypes.c_longlong * MAX_RANK)()
        _chk(self.lib.satt_h5_shape(self.ptr, name.encode(), ctypes.byref(r), d), f"shape {name}")
        return tuple(int(d[k]) for k in range(r.value))

    def get(self, name):
        """hf%get(name, value): the array in its Fortran shape."""
        out = np.zeros(self.shape(name), order="F")
        _chk(self.lib.satt_h5_get(self.ptr, name.encode(), _d(out)), f"get {name}")
        return out

    def write_iter_2d(self, it, gr_name, svel, periods, lon, lat, stlo=None, stla=None):
        sv = _f(svel)
        stlo, stla = _c(stlo), _c(stla)
        _chk(self.lib.satt_h5_write_iter_2d(
            self.ptr, int(it), gr_name.encode(), _d(stlo), _d(stla),
            0 if stlo is None else len(stlo), _d(_c(periods)), len(periods), _d(_c(lon)),
            len(lon), _d(_c(lat)), len(lat), _d(sv)), f"write_iter {it}")

    def write_model_3d_init(self, lon, lat, dep, vs3d, vp3d=None, rho3d=None):
        _chk(self.lib.satt_h5_write_model_3d_init(
            self.ptr, _d(_c(lon)), len(lon), _d(_c(lat)), len(lat), _d(_c(dep)), len(dep),
            _d(_f(vs3d)), _d(_f(vp3d)), _d(_f(rho3d))), "write_model_3d_init")

    def write_model_3d_iter(self, it, vs3d, vp3d=None, rho3d=None):
        nx, ny, nz = np.shape(vs3d)
        _chk(self.lib.satt_h5_write_model_3d_iter(self.ptr, int(it), nx, ny, nz, _d(_f(vs3d)),
                                                  _d(_f(vp3d)), _d(_f(rho3d))),
             f"write_model_3d_iter {it}")

    def write_gradient_3d(self, it, gradient_s):
        nk, nx, ny, nz = np.shape(gradient_s)
        _chk(self.lib.satt_h5_write_gradient_3d(self.ptr, int(it), nk, nx, ny, nz,
                                                _d(_f(gradient_s))), f"write_gradient {it}")

    def write_kdensity_3d(self, it, gr_name, ker_density):
        nk, nx, ny, nz = np.shape(ker_density)
        _chk(self.lib.satt_h5_write_kdensity_3d(self.ptr, int(it), gr_name.encode(), nk, nx,
                                                ny, nz, _d(_f(ker_density))),
             f"write_kdensity {it}")


def write_model_2d(path, svel, periods, lon, lat, stlo=None, stla=None):
    """att_acqui_2d%write_model (with stations) or write_target_model."""
    stlo, stla = _c(stlo), _c(stla)
    _chk(load().satt_h5_write_model_2d(
        path.encode(), _d(stlo), _d(stla), 0 if stlo is None else len(stlo), _d(_c(periods)),
        len(periods), _d(_c(lon)), len(lon), _d(_c(lat)), len(lat), _d(_f(svel))),
        f"write_model_2d {path}")


def read_model_3d(path, shape, name="/vs"):
    """The initial model of init_model_type 2 (src/model.f90:95-105)."""
    nx, ny, nz = shape
    out = np.zeros((nx, ny, nz), order="F")
    _chk(load().satt_h5_read_model_3d(path.encode(), name.encode(), nx, ny, nz, _d(out)),
         f"read_model_3d {path}")
    return out
