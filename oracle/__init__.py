"""CPU oracle for the compact-FD6 / MHD-RHS / RK4 hot path.

TEST INFRASTRUCTURE ONLY: the FP64 C restatement of the reference algorithm that the HIP
path is checked against.  Only tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg import this package; the product (``cssx``) never does.

Each wrapped C function cites the reference file:line it restates in oracle/*.c.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def build(force: bool = False) -> str:
    """Compile oracle/*.c with gcc (make); returns the library path."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s"] + (["-B"] if force else []) + ["-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        L = _lib
        L.orc_factors_d1.argtypes = [ctypes.c_int] * 4 + [_dp] * 3
        L.orc_factors_d2.argtypes = [ctypes.c_int] * 4 + [_dp] * 3
        L.orc_deriv_pencil.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_long, _dp, ctypes.c_long,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_compact_fd6_py.argtypes = [ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int,
                                         ctypes.c_int, _dp, _dp]
        L.orc_deriv_field.argtypes = [ctypes.c_int] * 5 + [_dp, _dp] + [ctypes.c_int] * 3 + [_dp, _dp] + \
            [ctypes.c_double, ctypes.c_double, _dp]
        _vp = ctypes.c_void_p
        L.orc_rhs.argtypes = [_vp, _vp, _vp, _dp, _dp, _dp]
        L.orc_bound.argtypes = [_vp, _vp, _vp, _dp]
        L.orc_rk4.argtypes = [_vp, _vp, _vp, _dp, ctypes.c_double, ctypes.c_int]
        L.orc_diffusion.argtypes = [_vp, _vp, _vp, _dp, _dp, _dp]
        L.orc_rk2_tvd.argtypes = [_vp, _vp, _vp, _dp, ctypes.c_double, ctypes.c_int]
        L.orc_rk3_tvd.argtypes = [_vp, _vp, _vp, _dp, ctypes.c_double, ctypes.c_int]
        L.orc_step_size.argtypes = [_vp, _vp, _vp, _dp, ctypes.c_double, _dp]
        L.orc_lap_cross.argtypes = [_vp, _vp, _vp, _dp, _dp]
    return _lib


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ghost4(b):
    """4 ghost values; a scalar fills slot 0 (the wall value), None gives zeros."""
    g = np.zeros(4)
    if b is None:
        return g
    if np.ndim(b) == 0:
        g[0] = float(b)
        return g
    return _f64(b)


def factors(order: int, n: int, lo_wall: bool, hi_wall: bool, iv: int):
    """gam1, gam2, gam3 tables (derivatives.f90:106-518)."""
    g = [np.zeros(n) for _ in range(3)]
    fn = lib().orc_factors_d1 if order == 1 else lib().orc_factors_d2
    fn(n, int(lo_wall), int(hi_wall), iv, *[_ptr(x) for x in g])
    return tuple(g)


def deriv_pencil(order, x, lo_wall, hi_wall, bdry, b1=None, b2=None, scale=1.0, hneu=None, phi_form=False):
    """One Dbyd*/D2byd*2 pencil (derivatives.f90:547-1684)."""
    x = _f64(x)
    n = x.size
    out = np.empty(n)
    b1 = _ghost4(b1)
    b2 = _ghost4(b2)
    if hneu is None:
        hneu = scale
    rc = lib().orc_deriv_pencil(order, n, _ptr(x), 1, _ptr(out), 1, int(lo_wall), int(hi_wall), bdry,
                                _ptr(b1), _ptr(b2), scale, hneu, int(phi_form))
    if rc:
        raise ValueError(f"orc_deriv_pencil rc={rc}")
    return out


def compact_fd6_py(dri, arr, b1, b2, ibc, dtype=0, darr=None):
    """CSS-Python derivatives.compact_fd6 restated (derivatives.py:32-185).

    Returns (result, darr_after) where darr_after is what the Python leaves in darr."""
    arr = _f64(arr)
    n = arr.size
    d = np.zeros(n) if darr is None else _f64(darr).copy()
    bb1 = None if np.ndim(b1) == 0 else _f64(b1)
    bb2 = None if np.ndim(b2) == 0 else _f64(b2)
    res = np.empty(n)
    rc = lib().orc_compact_fd6_py(float(dri), n, _ptr(arr), _ptr(bb1), _ptr(bb2), int(ibc), int(dtype),
                                  _ptr(d), _ptr(res))
    if rc:
        raise ValueError(f"orc_compact_fd6_py rc={rc}")
    return res, d


def deriv_field(axis, order, f, lo_wall=True, hi_wall=True, bdry=0, ghost_lo=None, ghost_hi=None,
                scale=1.0, hneu=None, metric=None):
    """Whole-subdomain Dbydr/Dbydt/Dbydp on f[nr, nphi, nth] (C order = reference layout)."""
    f = _f64(f)
    nr, nphi, nth = f.shape
    out = np.empty_like(f)
    gl = None if ghost_lo is None else _f64(ghost_lo)
    gh = None if ghost_hi is None else _f64(ghost_hi)
    m = None if metric is None else _f64(metric)
    if hneu is None:
        hneu = scale if order == 1 else np.sqrt(scale)
    rc = lib().orc_deriv_field(axis, order, nth, nphi, nr, _ptr(f), _ptr(out), int(lo_wall), int(hi_wall),
                               bdry, _ptr(gl), _ptr(gh), float(scale), float(hneu), _ptr(m))
    if rc:
        raise ValueError(f"orc_deriv_field rc={rc}")
    return out


def _structs(model):
    g = model.grid_struct()
    p = model.params_struct()
    pr = model.profiles_struct()
    return g, p, pr


def rhs(model, state, want_b=False):
    """RHS (rhs_mpi.f90:279-783) of a state (nv, nr, nphi, nth); returns dvars (and B)."""
    g, p, pr = _structs(model)
    st = _f64(state)
    dv = np.empty_like(st)
    B = np.empty((3,) + st.shape[1:]) if want_b else None
    rc = lib().orc_rhs(ctypes.addressof(g), ctypes.addressof(p), ctypes.addressof(pr), _ptr(st), _ptr(dv),
                       _ptr(B))
    if rc:
        raise ValueError(f"orc_rhs rc={rc}")
    return (dv, B) if want_b else dv


def bound(model, state):
    """Bound (rhs_mpi.f90:2395-2491) on a copy of the state."""
    g, p, pr = _structs(model)
    st = _f64(state).copy()
    lib().orc_bound(ctypes.addressof(g), ctypes.addressof(p), ctypes.addressof(pr), _ptr(st))
    return st


def diffusion(model, state, B=None):
    """Compute_Diffusion(1) (diffusion.f90:70-474) of a state (A in slots 6-8): the SLD
    tendency.  B = curl A as RHS(1) leaves it in vars1(6:8); computed when not given."""
    g, p, pr = _structs(model)
    st = _f64(state)
    nv = st.shape[0]
    v = st.copy()
    a = np.ascontiguousarray(st[5:8]) if nv == 8 else np.zeros((3,) + st.shape[1:])
    if nv == 8:
        if B is None:
            B = rhs(model, st, want_b=True)[1]
        v[5:8] = B
    dv = np.empty_like(st)
    rc = lib().orc_diffusion(ctypes.addressof(g), ctypes.addressof(p), ctypes.addressof(pr), _ptr(v), _ptr(a),
                             _ptr(dv))
    if rc:
        raise ValueError(f"orc_diffusion rc={rc}")
    return dv


def rk4(model, state, dt, nsteps, scheme="rk4"):
    """nsteps of Step_RK4 (step_rk4.f90:339-493), or of Step_RK3_TVD (:166-335) / Step_RK2_TVD
    (:31-162) for scheme 'rk3_tvd' / 'rk2_tvd', on a copy of the state."""
    g, p, pr = _structs(model)
    st = _f64(state).copy()
    fn = {"rk4": lib().orc_rk4, "rk3_tvd": lib().orc_rk3_tvd, "rk2_tvd": lib().orc_rk2_tvd}[scheme]
    rc = fn(ctypes.addressof(g), ctypes.addressof(p), ctypes.addressof(pr), _ptr(st), float(dt), int(nsteps))
    if rc:
        raise ValueError(f"{scheme} rc={rc}")
    return st


def step_size(model, state, cfl_safety=None):
    """Compute_Step_Size (rhs_mpi.f90:92-275) of an FP64 state (A in slots 6-8).

    Returns a dict: dt (with the safety factor), the crossing-time minima dtmaxc (sound and
    flow), dtmaxb (with the Alfven speed), dtmaxd (explicit-diffusion limit, 1e10 without
    Laplacian) and the conservation totals E, L, m, mrsint (:156-183)."""
    g, p, pr = _structs(model)
    st = _f64(state)
    out = np.zeros(8)
    cs = model.c.cfl_safety if cfl_safety is None else cfl_safety
    rc = lib().orc_step_size(ctypes.addressof(g), ctypes.addressof(p), ctypes.addressof(pr), _ptr(st), float(cs),
                             _ptr(out))
    if rc:
        raise ValueError(f"orc_step_size rc={rc}")
    return {"dt": out[0], "E": out[1], "L": out[2], "m": out[3], "mrsint": out[4], "dtmaxc": out[5],
            "dtmaxb": out[6], "dtmaxd": out[7]}


def lap_cross(model, state):
    """The six sweeps of Laplacian_Cross_Terms_Radial/_Theta (rhs_mpi.f90:2070-2296) that the
    RHS forms for a state (Laplacian on): d(r X_k)/dr for X = (dw/dth/r, du/dph/(r sin),
    du/dth/r, dv/dph/(r sin)), then d(sin X_4)/dth and d(sin dw/dph/(r sin))/dth."""
    g, p, pr = _structs(model)
    st = _f64(state)
    out = np.empty((6,) + st.shape[1:])
    rc = lib().orc_lap_cross(ctypes.addressof(g), ctypes.addressof(p), ctypes.addressof(pr), _ptr(st), _ptr(out))
    if rc:
        raise ValueError(f"orc_lap_cross rc={rc}")
    return out
