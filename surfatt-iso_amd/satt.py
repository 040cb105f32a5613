"""Python binding of libsatt.so (the C-ABI in include/satt.h) via ctypes.

The product is the HIP library; this module only marshals numpy arrays into
its entry points, for tests and measurement.  It never falls back to a CPU
path: when libsatt.so is missing or no GPU is present, every call raises.

Field convention: numpy arrays of shape (nx, ny) in Fortran order, so that
``f[i, j]`` is the reference's ``f(i+1, j+1)``; model arrays have shape
(nperiod, nx, ny) in Fortran order (period fastest), the reference's
att_grid layout (src/grid.f90:25-34).
"""
import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsatt.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "satt.h")

OK, W_MAXITER = 0, 1
E_ARG, E_SRC_ON_BOUNDARY, E_REC_OUT_OF_GRID, E_NOMEM, E_HIP, E_NODEV = -1, -2, -3, -4, -5, -6
E_TEAM_WAIT = -7
E_DISPERSION = -8
PREC_FP64, PREC_MIXED = 0, 1
LAUNCH_BATCHED, LAUNCH_PER_SOLVE = 0, 1
FIELD_T, FIELD_TA, FIELD_TD = 0, 1, 2
NOALLOC_MAX_CELLS, NOALLOC_MAX_AXIS = 262144, 8192   # satt_fwd_lonlat_noalloc's scratch

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None


def topo_sigma(vel, period, wavelen_factor):
    """sigma (degrees) of one period for the topographic smoothing,
    src/grid.f90:116."""
    return float(load().satt_topo_sigma(float(vel), float(period), float(wavelen_factor)))


class SattError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"status {status}: {msg}")
        self.status = status


class Batch(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("nperiod", ctypes.c_int), ("xgrids", _dp), ("ygrids", _dp),
        ("svel", _dp), ("m11", _dp), ("m12", _dp), ("m22", _dp),
        ("nitems", ctypes.c_int), ("item_pidx", _ip), ("item_srcx", _dp),
        ("item_srcy", _dp), ("rec_off", _ip), ("recx", _dp), ("recy", _dp),
        ("tt", _dp), ("weight", _dp), ("idx_path", _ip), ("npath", ctypes.c_int),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("tt_fwd", _dp), ("chi", _dp), ("adj_s", _dp), ("adj_density", _dp),
        ("item_chi", _dp), ("iters", _ip), ("item_status", _ip), ("weight_out", _dp),
        ("iters_fp32", _ip),
    ]


class PlanStats(ctypes.Structure):
    _fields_ = [
        ("chunk", ctypes.c_int), ("launches", ctypes.c_int), ("device_bytes", ctypes.c_longlong),
        ("ms_forward", ctypes.c_double), ("ms_adjoint", ctypes.c_double),
        ("ms_total", ctypes.c_double), ("fwd_sweeps", ctypes.c_longlong),
        ("adj_sweeps", ctypes.c_longlong), ("fwd_sweeps32", ctypes.c_longlong),
        ("adj_sweeps32", ctypes.c_longlong), ("ms_fwd_fp64", ctypes.c_double),
        ("ms_fwd_fp32", ctypes.c_double), ("ms_adj_fp64", ctypes.c_double),
        ("ms_adj_fp32", ctypes.c_double), ("ms_fwd_setup", ctypes.c_double),
        ("ms_adj_setup", ctypes.c_double),
    ]


def header_symbols():
    """Every function name declared in include/satt.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(satt_\w+)\s*\(", text, re.M)))


def load(path=LIB_PATH):
    """Load libsatt.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SattError(E_ARG, f"{path} not built (run __graft_entry__.build())")
    L = ctypes.CDLL(path)
    for name in header_symbols():
        fn = getattr(L, name)
        fn.restype = ctypes.c_int
    L.satt_destroy.restype = None
    L.satt_plan_destroy.restype = None
    L.satt_forward_simulate_release.restype = None
    L.satt_server_destroy.restype = None
    L.satt_last_error.restype = ctypes.c_char_p
    L.satt_version.restype = ctypes.c_char_p
    L.satt_topo_sigma.restype = ctypes.c_double
    L.satt_topo_sigma.argtypes = [ctypes.c_double] * 3
    L.satt_last_device_ms.restype = ctypes.c_double
    _lib = L
    return L


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _i(a):
    return None if a is None else a.ctypes.data_as(_ip)


def _f(a, shape=None):
    a = np.asfortranarray(np.asarray(a, dtype=np.float64))
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


class Context:
    """A satt_ctx: one HIP stream on one device."""

    def __init__(self, device=0):
        self.lib = load()
        p = ctypes.c_void_p()
        st = self.lib.satt_create(ctypes.byref(p), int(device))
        if st != OK:
            raise SattError(st, "satt_create failed (no usable GPU?)")
        self.ptr = p

    def tune_waves(self, nx, ny, nitems, precision=PREC_MIXED):
        """Autotuner of the launch shape (team size and prefetch depth per
        sweep stage) for an nx x ny grid and a launch of nitems items; returns
        {"waves": W, "ms": [ms for W=1..8], "stage_waves": [W per sweep stage],
        "stage_ms": [[ms per stage] for W=1..8], "stage_depth": [2 or 4 per
        stage], "stage_ms_deep": [[ms per stage] at the chosen W, at W=1]} and
        keeps the choice in the context's table."""
        r = TuneResult()
        self._check(self.lib.satt_tune_waves(self.ptr, int(nx), int(ny), int(nitems),
                                             int(precision), ctypes.byref(r)))
        return r.as_dict()

    def close(self):
        if self.ptr:
            self.lib.satt_destroy(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def error(self):
        return self.lib.satt_last_error(self.ptr).decode()

    def _check(self, st, allow_warn=True):
        if st < 0 or (st > 0 and not allow_warn):
            raise SattError(st, self.error())
        return st

    # ---- per-solve entry points
    def fwd(self, xx, yy, spha, sphb, sphc, fun, x0, y0):
        """satt_fwd_lonlat -> (T, iters, status)."""
        xx, yy = _c(xx), _c(yy)
        sh = (len(xx), len(yy))
        spha, sphb, sphc, fun = (_f(v, sh) for v in (spha, sphb, sphc, fun))
        T = np.zeros(sh, order="F")
        it = ctypes.c_int(0)
        st = self.lib.satt_fwd_lonlat(self.ptr, _d(xx), _d(yy), sh[0], sh[1], _d(spha), _d(sphb),
                                      _d(sphc), _d(fun), ctypes.c_double(x0), ctypes.c_double(y0),
                                      _d(T), ctypes.byref(it))
        self._check(st)
        return T, it.value, st

    def fwd_noalloc(self, xx, yy, spha, sphb, sphc, fun, x0, y0):
        """satt_fwd_lonlat_noalloc (no device allocation) -> (T, iters, status)."""
        xx, yy = _c(xx), _c(yy)
        sh = (len(xx), len(yy))
        spha, sphb, sphc, fun = (_f(v, sh) for v in (spha, sphb, sphc, fun))
        T = np.zeros(sh, order="F")
        it = ctypes.c_int(0)
        st = self.lib.satt_fwd_lonlat_noalloc(
            self.ptr, _d(xx), _d(yy), sh[0], sh[1], _d(spha), _d(sphb), _d(sphc), _d(fun),
            ctypes.c_double(x0), ctypes.c_double(y0), _d(T), ctypes.byref(it))
        self._check(st)
        return T, it.value, st

    def adj(self, xx, yy, spha, sphb, sphc, T, xrec, yrec, src_adj):
        """satt_adj_lonlat -> (Ta, iters, status)."""
        xx, yy = _c(xx), _c(yy)
        sh = (len(xx), len(yy))
        spha, sphb, sphc, T = (_f(v, sh) for v in (spha, sphb, sphc, T))
        xrec, yrec, src_adj = _c(xrec), _c(yrec), _c(src_adj)
        Ta = np.zeros(sh, order="F")
        it = ctypes.c_int(0)
        st = self.lib.satt_adj_lonlat(self.ptr, _d(xx), _d(yy), sh[0], sh[1], _d(spha), _d(sphb),
                                      _d(sphc), _d(T), _d(xrec), _d(yrec), _d(src_adj),
                                      len(xrec), _d(Ta), ctypes.byref(it))
        self._check(st)
        return Ta, it.value, st

    def mask(self, xx, yy, table, x0, y0):
        xx, yy = _c(xx), _c(yy)
        t = _f(table, (len(xx), len(yy))).copy(order="F")
        self._check(self.lib.satt_mask(self.ptr, _d(xx), _d(yy), len(xx), len(yy), _d(t),
                                       ctypes.c_double(x0), ctypes.c_double(y0)))
        return t

    def interp2(self, x, y, v, xq, yq):
        x, y = _c(x), _c(y)
        v = _f(v, (len(x), len(y)))
        xq, yq = _c(xq), _c(yq)
        out = np.zeros(len(xq))
        self._check(self.lib.satt_interp2(self.ptr, _d(x), _d(y), len(x), len(y), _d(v), _d(xq),
                                          _d(yq), len(xq), _d(out)))
        return out

    # ---- model-side steps (src/shared/utils.f90, src/grid.f90, src/tomo2d.f90)
    @staticmethod
    def _pm(v, nf, sh):
        v = np.asarray(v, dtype=np.float64)
        if v.shape == sh and nf == 1:
            v = v.reshape((1,) + sh)
        if v.shape != (nf,) + sh:
            raise SattError(E_ARG, f"expected shape {(nf,) + sh}, got {v.shape}")
        return np.asfortranarray(v)

    def smooth_geo2(self, lons, lats, data, sigma):
        """gaussian_smooth_geo_2 (src/shared/utils.f90:1194-1223) of each field of
        ``data`` (nx, ny) or (nfield, nx, ny) with ``sigma`` (scalar or per field,
        degrees); a single field with several sigmas is smoothed once per sigma.
        Returns (nfield, nx, ny), or (nx, ny) for one field and one sigma."""
        lons, lats = _c(lons), _c(lats)
        sh = (len(lons), len(lats))
        sig = np.atleast_1d(np.asarray(sigma, dtype=np.float64)).copy()
        data = np.asarray(data, dtype=np.float64)
        ndata = 1 if data.shape == sh else data.shape[0]
        if sig.size == 1 and ndata > 1:
            sig = np.full(ndata, sig[0])
        nf = sig.size
        src = self._pm(data, ndata, sh)
        out = np.zeros((nf,) + sh, order="F")
        self._check(self.lib.satt_smooth_geo2(self.ptr, _d(lons), _d(lats), sh[0], sh[1], nf,
                                              ndata, _d(src), _d(sig), _d(out)))
        return out[0] if (nf == 1 and data.shape == sh) else out

    def gradient_geo2(self, lons, lats, f):
        """gradient_2_geo (src/shared/utils.f90:1118-1160): (tx, ty) of ``f``,
        (nx, ny) or (nfield, nx, ny)."""
        lons, lats = _c(lons), _c(lats)
        sh = (len(lons), len(lats))
        f = np.asarray(f, dtype=np.float64)
        one = f.shape == sh
        nf = 1 if one else f.shape[0]
        src = self._pm(f, nf, sh)
        tx, ty = np.zeros((nf,) + sh, order="F"), np.zeros((nf,) + sh, order="F")
        self._check(self.lib.satt_gradient_geo2(self.ptr, _d(lons), _d(lats), sh[0], sh[1], nf,
                                                _d(src), _d(tx), _d(ty)))
        return (tx[0], ty[0]) if one else (tx, ty)

    def topo_metric(self, lons, lats, z, sigma, angle=True):
        """The metric of att_grid%get_topo (src/grid.f90:114-123,137-141):
        (m11, m12, m22[, angle]) of shape (nperiod, nx, ny) for topography ``z``
        (nx, ny, km) and per-period ``sigma`` (degrees)."""
        lons, lats = _c(lons), _c(lats)
        sh = (len(lons), len(lats))
        z = _f(z, sh)
        sig = _c(np.atleast_1d(sigma))
        nf = len(sig)
        outs = [np.zeros((nf,) + sh, order="F") for _ in range(4 if angle else 3)]
        self._check(self.lib.satt_topo_metric(self.ptr, _d(lons), _d(lats), sh[0], sh[1], _d(z),
                                              nf, _d(sig), _d(outs[0]), _d(outs[1]),
                                              _d(outs[2]), _d(outs[3]) if angle else None))
        return tuple(outs)

    def steepest_descent(self, lons, lats, adj_s, sigma_2d, updatemax, svel):
        """att_tomo_2d%optimize + steepest_descent (src/tomo2d.f90:203-251):
        returns (svel_new, ker_s), both (nperiod, nx, ny)."""
        lons, lats = _c(lons), _c(lats)
        sh = (len(lons), len(lats))
        adj_s = np.asfortranarray(adj_s, dtype=np.float64)
        nf = adj_s.shape[0]
        sv = self._pm(svel, nf, sh).copy(order="F")
        ker = np.zeros((nf,) + sh, order="F")
        self._check(self.lib.satt_steepest_descent(
            self.ptr, _d(lons), _d(lats), sh[0], sh[1], nf, _d(self._pm(adj_s, nf, sh)),
            ctypes.c_double(sigma_2d), ctypes.c_double(updatemax), _d(sv), _d(ker)))
        return sv, ker

    def combine_kernels(self, adj_s, adj_density, sen_vs, sen_vp, sen_rho, vs3d=None,
                        vp3d=None, use_alpha_beta_rho=False, rho_scaling=True, ker=None,
                        ker_density=None):
        """att_acqui%combine_kernels (src/acqui.f90:197-245): (ker,
        ker_density), each (nker, nx, ny, nz), accumulated into the given
        arrays (else zeros)."""
        sv = _f(sen_vs)
        npd, nx, ny, nz = sv.shape
        nker = 3 if use_alpha_beta_rho else 1
        k = np.zeros((nker, nx, ny, nz), order="F") if ker is None else _f(ker).copy(order="F")
        kd = (np.zeros((nker, nx, ny, nz), order="F") if ker_density is None
              else _f(ker_density).copy(order="F"))
        opt = lambda a: None if a is None else _f(a)
        args = [_f(adj_s), _f(adj_density), sv, _f(sen_vp), opt(sen_rho), opt(vs3d), opt(vp3d)]
        self._check(self.lib.satt_combine_kernels(
            self.ptr, npd, nx, ny, nz, *(_d(a) for a in args), int(bool(use_alpha_beta_rho)),
            int(bool(rho_scaling)), _d(k), _d(kd)))
        return k, kd

    def correct_depth(self, ker_vs, ker_vp, ker_rho, dip_angle, depz):
        """correct_depth (src/surfker/surfker.f90:260-283): corrected copies of
        the (nperiod, nx, ny, nz) depth kernels for dip angles (nperiod, nx, ny)."""
        ks = [_f(k).copy(order="F") for k in (ker_vs, ker_vp, ker_rho)]
        npd, nx, ny, nz = ks[0].shape
        self._check(self.lib.satt_correct_depth(
            self.ptr, npd, nx, ny, nz, _d(ks[0]), _d(ks[1]), _d(ks[2]), _d(_f(dip_angle)),
            _d(_c(depz))))
        return tuple(ks)

    @staticmethod
    def _inv(xinv, yinv, zinv):
        xinv, yinv, zinv = _f(xinv), _f(yinv), _f(zinv)
        if xinv.ndim != 2 or yinv.shape[1:] != xinv.shape[1:] or zinv.shape[1:] != xinv.shape[1:]:
            raise SattError(E_ARG, "inversion grids are (n, nset) arrays with one nset")
        return xinv, yinv, zinv, xinv.shape[0], yinv.shape[0], zinv.shape[0], xinv.shape[1]

    def inv_grid_iso(self, xinv, yinv, zinv, adj, vel_x, vel_y, vel_z):
        """inv_grid_iso (src/eikon_solver_2d.f90:1538): gk (nset*nxinv*nyinv*nzinv)."""
        xinv, yinv, zinv, nxi, nyi, nzi, ns = self._inv(xinv, yinv, zinv)
        adj = _f(adj)
        nx, ny, nz = adj.shape
        gk = np.zeros(ns * nxi * nyi * nzi)
        self._check(self.lib.satt_inv_grid_iso(
            self.ptr, _d(xinv), _d(yinv), _d(zinv), nxi, nyi, nzi, ns, _d(adj), nx, ny, nz,
            _d(gk), _d(_c(vel_x)), _d(_c(vel_y)), _d(_c(vel_z))))
        return gk

    def inv2fwd_iso(self, xinv, yinv, zinv, gk, vel_x, vel_y, vel_z):
        """inv2fwd_iso (src/eikon_solver_2d.f90:1613): update (nx, ny, nz)."""
        xinv, yinv, zinv, nxi, nyi, nzi, ns = self._inv(xinv, yinv, zinv)
        vx, vy, vz = _c(vel_x), _c(vel_y), _c(vel_z)
        up = np.zeros((len(vx), len(vy), len(vz)), order="F")
        self._check(self.lib.satt_inv2fwd_iso(
            self.ptr, _d(xinv), _d(yinv), _d(zinv), nxi, nyi, nzi, ns, len(vx), len(vy), len(vz),
            _d(_c(gk)), _d(vx), _d(vy), _d(vz), _d(up)))
        return up

    def kernel_postproc(self, vel_x, vel_y, vel_z, xinv, yinv, zinv, ker, ker_density,
                        kdensity_coe, chi0, weight, gradient_s):
        """att_acqui%post_processing_for_kernel (src/acqui.f90:247-285): returns
        (ker after the density preconditioning, gradient_s after the update)."""
        xinv, yinv, zinv, nxi, nyi, nzi, ns = self._inv(xinv, yinv, zinv)
        k, g = _f(ker).copy(order="F"), _f(gradient_s).copy(order="F")
        nker, nx, ny, nz = k.shape
        kd = None if ker_density is None else _f(ker_density)
        self._check(self.lib.satt_kernel_postproc(
            self.ptr, nker, nx, ny, nz, _d(_c(vel_x)), _d(_c(vel_y)), _d(_c(vel_z)), ns, nxi,
            nyi, nzi, _d(xinv), _d(yinv), _d(zinv), _d(k), _d(kd), ctypes.c_double(kdensity_coe),
            ctypes.c_double(chi0), ctypes.c_double(weight), _d(g)))
        return k, g

    def fwdsurf3d(self, vs3d, depz, periods, iwave=2, igr=0, status=False):
        """Surface-wave velocity maps of att_grid%fwdsurf (src/grid.f90:150-163):
        (nperiod, nx, ny) phase (igr 0) or group (igr > 0) velocities of Love
        (iwave 1) or Rayleigh (iwave 2) waves for vs3d (nx, ny, nz) on depths
        depz (nz).  status=True also returns the per-column codes and does not
        raise on columns without a root."""
        vs = np.asfortranarray(vs3d, dtype=np.float64)
        nx, ny, nz = vs.shape
        depz, periods = _c(depz), _c(periods)
        out = np.zeros((len(periods), nx, ny), order="F")
        cs = np.zeros((nx, ny), dtype=np.int32, order="F")
        st = self.lib.satt_fwdsurf3d(self.ptr, _d(vs), nx, ny, nz, _d(depz), len(periods),
                                     _d(periods), int(iwave), int(igr), _d(out), _i(cs))
        if status and st == E_DISPERSION:
            return out, cs
        self._check(st)
        return (out, cs) if status else out

    def depthkernel(self, vs3d, depz, periods, iwave=2, igr=0, status=False):
        """satt_depthkernel: depthkernel_decomp's sen_vs, sen_vp, sen_rho
        (nperiod, nx, ny, nz) and the phase velocities cp (nperiod, nx, ny)
        of vs3d (nx, ny, nz); with status also the per-column codes."""
        vs = _f(vs3d)
        nx, ny, nz = vs.shape
        depz, per = _c(depz), _c(periods)
        out = [np.zeros((len(per), nx, ny, nz), order="F") for _ in range(3)]
        cp = np.zeros((len(per), nx, ny), order="F")
        st = np.zeros(nx * ny, dtype=np.int32)
        rc = self.lib.satt_depthkernel(self.ptr, _d(vs), nx, ny, nz, _d(depz), len(per), _d(per),
                                       int(iwave), int(igr), *(_d(o) for o in out), _d(cp), _i(st))
        if status:
            return (*out, cp, st.reshape((nx, ny), order="F"), rc)
        self._check(rc)
        return (*out, cp)

    def hbm_probe(self, nbytes=1 << 30):
        """(read GB/s, copy GB/s) of streaming kernels over nbytes."""
        r, c = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.satt_hbm_probe(self.ptr, ctypes.c_double(nbytes), ctypes.byref(r),
                                            ctypes.byref(c)))
        return r.value, c.value

    def lds_probe(self):
        """(LDS read GB/s, LDS peak GB/s at 256 bytes per clock per CU)."""
        r, pk = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.satt_lds_probe(self.ptr, ctypes.byref(r), ctypes.byref(pk)))
        return r.value, pk.value

    def last_device_ms(self):
        """Device time of the last model-side call's kernels (ms)."""
        return float(self.lib.satt_last_device_ms(self.ptr))

    def fwd_setup_fields(self, xx, yy, spha, sphb, sphc, fun, x0, y0):
        xx, yy = _c(xx), _c(yy)
        sh = (len(xx), len(yy))
        spha, sphb, sphc, fun = (_f(v, sh) for v in (spha, sphb, sphc, fun))
        out = [np.zeros(sh, order="F") for _ in range(6)]
        self._check(self.lib.satt_fwd_setup_fields(
            self.ptr, _d(xx), _d(yy), sh[0], sh[1], _d(spha), _d(sphb), _d(sphc), _d(fun),
            ctypes.c_double(x0), ctypes.c_double(y0), *(_d(o) for o in out)))
        return tuple(out)

    def adj_setup_fields(self, xx, yy, spha, sphb, sphc, xrec, yrec, src_adj):
        xx, yy = _c(xx), _c(yy)
        sh = (len(xx), len(yy))
        spha, sphb, sphc = (_f(v, sh) for v in (spha, sphb, sphc))
        xrec, yrec, src_adj = _c(xrec), _c(yrec), _c(src_adj)
        out = [np.zeros(sh, order="F") for _ in range(4)]
        self._check(self.lib.satt_adj_setup_fields(
            self.ptr, _d(xx), _d(yy), sh[0], sh[1], _d(spha), _d(sphb), _d(sphc), _d(xrec),
            _d(yrec), _d(src_adj), len(xrec), *(_d(o) for o in out)))
        return tuple(out)

    # ---- batched
    def plan(self, batch, precision=PREC_FP64):
        return Plan(self, batch, precision)

    def plan_multi(self, batches, precision=PREC_FP64):
        """One plan over several grids (one batch dictionary per grid)."""
        return Plan(self, list(batches), precision)

    def forward_simulate(self, batch, istotable=True, isadj=True, precision=PREC_FP64):
        with Plan(self, batch, precision) as p:
            p.run(istotable, isadj)
            return p.results()

    def forward_simulate_dropin(self, batch, istotable=True, isadj=True, precision=PREC_FP64):
        """satt_forward_simulate, the drop-in call (its plan stays resident in
        the context for the next call with the same gathers)."""
        out, r = _result_arrays(batch)
        st = self.lib.satt_forward_simulate(self.ptr, ctypes.byref(_batch_struct(batch)),
                                            int(istotable), int(isadj), int(precision),
                                            ctypes.byref(r))
        self._check(st)
        out["status"] = st
        out["chi"] = float(out["chi"][0])
        return out

    def forward_simulate_release(self):
        self.lib.satt_forward_simulate_release(self.ptr)


def _result_arrays(b):
    """Output arrays of one part's results and the satt_result pointing at them."""
    nit, npd, nx, ny = len(b["item_pidx"]), b["nperiod"], b["nx"], b["ny"]
    out = {
        "tt_fwd": np.zeros(b["npath"]),
        "chi": np.zeros(1),
        "adj_s": np.zeros((npd, nx, ny), order="F"),
        "adj_density": np.zeros((npd, nx, ny), order="F") if b["dim"] == 3 else None,
        "item_chi": np.zeros(nit),
        "iters": np.zeros((nit, 3), dtype=np.int32),
        "iters_fp32": np.zeros((nit, 3), dtype=np.int32),
        "item_status": np.zeros(nit, dtype=np.int32),
        "weight_out": np.zeros(len(b["recx"])),
    }
    r = Result()
    r.tt_fwd, r.chi, r.adj_s = _d(out["tt_fwd"]), _d(out["chi"]), _d(out["adj_s"])
    r.adj_density = _d(out["adj_density"])
    r.item_chi, r.iters, r.item_status = _d(out["item_chi"]), _i(out["iters"]), _i(out["item_status"])
    r.weight_out = _d(out["weight_out"])
    r.iters_fp32 = _i(out["iters_fp32"])
    return out, r


def _batch_struct(b):
    s = Batch()
    s.dim, s.nx, s.ny, s.nperiod = b["dim"], b["nx"], b["ny"], b["nperiod"]
    s.xgrids, s.ygrids = _d(b["xgrids"]), _d(b["ygrids"])
    s.svel, s.m11, s.m12, s.m22 = _d(b["svel"]), _d(b["m11"]), _d(b["m12"]), _d(b["m22"])
    s.nitems = len(b["item_pidx"])
    s.item_pidx, s.item_srcx, s.item_srcy = _i(b["item_pidx"]), _d(b["item_srcx"]), _d(b["item_srcy"])
    s.rec_off, s.recx, s.recy = _i(b["rec_off"]), _d(b["recx"]), _d(b["recy"])
    s.tt, s.weight, s.idx_path, s.npath = _d(b["tt"]), _d(b["weight"]), _i(b["idx_path"]), b["npath"]
    return s


class Plan:
    """satt_plan: model, gathers and fields resident on the device.  `batch`
    is one batch dictionary, or a list of them (one grid each)."""

    def __init__(self, ctx, batch, precision=PREC_FP64):
        self.ctx, self.lib = ctx, ctx.lib
        self.parts = batch if isinstance(batch, list) else [batch]
        self.b = self.parts[0]
        arr = (Batch * len(self.parts))(*[_batch_struct(x) for x in self.parts])
        self._bs = arr
        p = ctypes.c_void_p()
        st = self.lib.satt_plan_create_multi(ctx.ptr, len(self.parts), arr, int(precision),
                                             ctypes.byref(p))
        ctx._check(st)
        self.ptr = p

    def close(self):
        if self.ptr:
            self.lib.satt_plan_destroy(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_model(self, svel, m11, m12, m22, part=0):
        arrs = [_f(v) for v in (svel, m11, m12, m22)]
        self.ctx._check(self.lib.satt_plan_set_model_part(self.ptr, int(part),
                                                          *(_d(a) for a in arrs)))

    def linesearch(self, svel, ker_s, chi0, updatemax, maxshrink, max_sub_niter):
        """att_tomo_2d%linesearch (src/tomo2d.f90:253-288) on this one-grid
        plan: returns {"svel": the last model tried, "updatemax": the step
        after the search, "chi": the last misfit, "nsub": sub-iterations run,
        "status"}.  The plan keeps the returned model."""
        sv, ker = _f(svel), _f(ker_s)
        out = np.zeros(sv.shape, order="F")
        upd, chi, nsub = ctypes.c_double(updatemax), ctypes.c_double(0), ctypes.c_int(0)
        st = self.ctx._check(self.lib.satt_plan_linesearch(
            self.ptr, _d(sv), _d(ker), ctypes.c_double(chi0), ctypes.c_double(maxshrink),
            int(max_sub_niter), ctypes.byref(upd), _d(out), ctypes.byref(chi),
            ctypes.byref(nsub)))
        return {"svel": out, "updatemax": upd.value, "chi": chi.value, "nsub": nsub.value,
                "status": st}

    def run(self, istotable=True, isadj=True, launch=LAUNCH_BATCHED):
        self.flags = (istotable, isadj)
        self.ctx._check(self.lib.satt_plan_run(self.ptr, int(istotable), int(isadj), int(launch)))

    def set_waves(self, waves):
        """Waves per item of the sweep kernels (capped per launch group)."""
        self.ctx._check(self.lib.satt_plan_set_waves(self.ptr, int(waves)))

    def set_stage_waves(self, waves):
        """Team size per sweep stage (forward FP32, forward FP64, adjoint FP32,
        adjoint FP64)."""
        w = np.ascontiguousarray(waves, dtype=np.int32)
        if w.shape != (4,):
            raise SattError(E_ARG, "four stage team sizes expected")
        self.ctx._check(self.lib.satt_plan_set_stage_waves(self.ptr, _i(w)))

    def set_stage_depth(self, depth):
        """Prefetch depth (2 or 4 steps) per sweep stage, in the order of
        set_stage_waves."""
        d = np.ascontiguousarray(depth, dtype=np.int32)
        if d.shape != (4,):
            raise SattError(E_ARG, "four stage depths expected")
        self.ctx._check(self.lib.satt_plan_set_stage_depth(self.ptr, _i(d)))

    def tune_waves(self):
        """Autotune the team size and prefetch depth of each sweep stage on
        this plan's inputs (as Context.tune_waves)."""
        r = TuneResult()
        self.ctx._check(self.lib.satt_plan_tune_waves(self.ptr, ctypes.byref(r)))
        return r.as_dict()

    def set_graph(self, on=True):
        """Replay form: batched runs capture their launch chain into a HIP
        graph once per (istotable, isadj) and replay it afterwards."""
        self.ctx._check(self.lib.satt_plan_set_graph(self.ptr, int(bool(on))))

    def wait(self):
        self.ctx._check(self.lib.satt_plan_wait(self.ptr))

    def results(self, part=0):
        out, r = _result_arrays(self.parts[part])
        st = self.lib.satt_plan_results_part(self.ptr, int(part), ctypes.byref(r))
        out["status"] = st
        out["chi"] = float(out["chi"][0])
        if st < 0:
            raise SattError(st, self.ctx.error())
        return out

    def item_field(self, item, which=FIELD_T):
        """Natural-layout field of one item of the last run's last launch group."""
        out = np.zeros((self.b["nx"], self.b["ny"]), order="F")
        self.ctx._check(self.lib.satt_plan_item_field(self.ptr, int(item), int(which), _d(out)))
        return out

    def stats(self):
        s = PlanStats()
        self.ctx._check(self.lib.satt_plan_get_stats(self.ptr, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in PlanStats._fields_}


# ------------------------------------------------------------------ serving

class TuneResult(ctypes.Structure):
    _fields_ = [("waves", ctypes.c_int), ("ms", ctypes.c_double * 8),
                ("stage_waves", ctypes.c_int * 4), ("stage_ms", (ctypes.c_double * 4) * 8),
                ("stage_depth", ctypes.c_int * 4), ("stage_ms_deep", (ctypes.c_double * 4) * 2)]

    def as_dict(self):
        return {"waves": self.waves, "ms": [float(v) for v in self.ms],
                "stage_waves": [int(v) for v in self.stage_waves],
                "stage_ms": [[float(v) for v in row] for row in self.stage_ms],
                "stage_depth": [int(v) for v in self.stage_depth],
                "stage_ms_deep": [[float(v) for v in row] for row in self.stage_ms_deep]}


class ServerConfig(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("max_batch", ctypes.c_int),
                ("streams", ctypes.c_int), ("device_bytes", ctypes.c_longlong),
                ("waves", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("xgrids", _dp), ("ygrids", _dp),
                ("svel", _dp), ("m11", _dp), ("m12", _dp), ("m22", _dp),
                ("srcx", ctypes.c_double), ("srcy", ctypes.c_double), ("nrec", ctypes.c_int),
                ("recx", _dp), ("recy", _dp), ("tt", _dp), ("weight", _dp)]


class Response(ctypes.Structure):
    _fields_ = [("tt_rec", _dp), ("kernel", _dp), ("T", _dp), ("chi", ctypes.c_double),
                ("status", ctypes.c_int), ("iters", ctypes.c_int * 2),
                ("iters_fp32", ctypes.c_int * 2), ("queue_ms", ctypes.c_double),
                ("latency_ms", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_longlong), ("completed", ctypes.c_longlong),
                ("launches", ctypes.c_longlong), ("max_in_flight", ctypes.c_longlong),
                ("device_busy_ms", ctypes.c_double)]


class Server:
    """satt_server: a request loop of independent fwd+adj solves.  A request
    is a dict with nx, ny, xgrids, ygrids, svel (nx, ny), optional m11, m12,
    m22, srcx, srcy and receiver arrays recx, recy, tt, weight."""

    def __init__(self, ctx, precision=PREC_FP64, max_batch=0, streams=0, device_bytes=0,
                 waves=0):
        self.ctx, self.lib = ctx, ctx.lib
        cfg = ServerConfig(int(precision), int(max_batch), int(streams), int(device_bytes),
                           int(waves))
        p = ctypes.c_void_p()
        ctx._check(self.lib.satt_server_create(ctx.ptr, ctypes.byref(cfg), ctypes.byref(p)))
        self.ptr = p
        self._live = {}

    def submit(self, req, want_T=False):
        nx, ny = int(req["nx"]), int(req["ny"])
        keep = {k: (_f(req[k]) if k in ("svel", "m11", "m12", "m22")
                    else np.ascontiguousarray(req[k], dtype=np.float64))
                for k in ("xgrids", "ygrids", "svel", "m11", "m12", "m22", "recx", "recy", "tt",
                          "weight") if req.get(k) is not None}
        nr = len(keep["recx"])
        q = Request()
        q.nx, q.ny, q.srcx, q.srcy, q.nrec = nx, ny, float(req["srcx"]), float(req["srcy"]), nr
        for k, v in keep.items():
            setattr(q, k, _d(v))
        out = {"tt_rec": np.zeros(nr), "kernel": np.zeros((nx, ny), order="F"),
               "T": np.zeros((nx, ny), order="F") if want_T else None}
        r = Response()
        r.tt_rec, r.kernel = _d(out["tt_rec"]), _d(out["kernel"])
        r.T = _d(out["T"]) if want_T else None
        t = ctypes.c_longlong(0)
        self.ctx._check(self.lib.satt_server_submit(self.ptr, ctypes.byref(q), ctypes.byref(r),
                                                    ctypes.byref(t)))
        self._live[t.value] = (r, out, keep)
        return t.value

    def test(self, ticket):
        return self.lib.satt_server_test(self.ptr, ctypes.c_longlong(ticket)) == 1

    def wait(self, ticket):
        st = self.lib.satt_server_wait(self.ptr, ctypes.c_longlong(ticket))
        r, out, _ = self._live.pop(ticket)
        out.update(status=r.status, chi=r.chi, iters=tuple(r.iters),
                   iters_fp32=tuple(r.iters_fp32), queue_ms=r.queue_ms, latency_ms=r.latency_ms)
        if st < 0:
            raise SattError(st, "request failed")
        return out

    def drain(self):
        self.ctx._check(self.lib.satt_server_drain(self.ptr))

    def stats(self):
        s = ServerStats()
        self.ctx._check(self.lib.satt_server_get_stats(self.ptr, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self.ptr:
            self.lib.satt_server_destroy(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
