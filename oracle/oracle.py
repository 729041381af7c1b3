"""ctypes loader for the CPU oracle (test infrastructure only).

Used by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py as the checker / timed CPU baseline.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_sw2.so")


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        _setup(_lib)
    return _lib


PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)


class ODesc(C.Structure):
    _fields_ = [("nnodes", C.c_int), ("nelems2d", C.c_int), ("nelems1d", C.c_int), ("nmat", C.c_int),
                ("x", PD), ("y", PD), ("z", PD), ("node_bc", PI), ("node_bc_val", PD),
                ("e2d", PI), ("e2d_mat", PI), ("e2d_ks", PD), ("e2d_src_flag", PI), ("e2d_src", PD),
                ("e1d", PI), ("e1d_adj", PI), ("e1d_bc", PI), ("e1d_val", PD),
                ("mat_turb", PI), ("mat_eev_mode", PI), ("mat_eev_coef", PD), ("mat_ev", PD),
                ("mat_cor_speed", PD),
                ("g", C.c_double), ("dt", C.c_double), ("tau_temporal", C.c_double), ("tau_pg", C.c_double),
                ("drying_lower", C.c_double), ("drying_upper", C.c_double), ("viscosity", C.c_double),
                ("wind", C.c_int), ("mat_windatt", PD), ("density", C.c_double)]


class OOpts(C.Structure):
    _fields_ = [("tol_nonlin", C.c_double), ("inc_nonlin", C.c_double), ("max_nonlin_it", C.c_int),
                ("max_lin_it", C.c_int), ("max_cuts", C.c_int), ("prec", C.c_int), ("nblock", C.c_int),
                ("lin_rtol", C.c_double), ("lin_atol", C.c_double), ("hist_cap", C.c_int)]


class OReport(C.Structure):
    _fields_ = [("converged", C.c_int), ("iterations", C.c_int), ("lin_iterations_total", C.c_int),
                ("solv_fail", C.c_int), ("nhist", C.c_int), ("hist_max", PD), ("hist_l2", PD),
                ("hist_inc", PD), ("hist_lin_it", PI), ("time_load", C.c_double), ("time_setup", C.c_double),
                ("time_solve", C.c_double), ("time_resid", C.c_double)]


def _setup(L):
    L.o_create.restype = C.c_void_p
    L.o_create.argtypes = [C.POINTER(ODesc)]
    L.o_destroy.argtypes = [C.c_void_p]
    for name in ["o_get_geometry"]:
        getattr(L, name).argtypes = [C.c_void_p, PD, PD, PD, PD]
    L.o_pattern_nnz.argtypes = [C.c_void_p]
    L.o_pattern_nnz.restype = C.c_int
    L.o_get_pattern.argtypes = [C.c_void_p, PI, PI]
    L.o_set_state.argtypes = [C.c_void_p] + [PD] * 6
    L.o_get_state.argtypes = [C.c_void_p] + [PD] * 6
    L.o_set_dt.argtypes = [C.c_void_p, C.c_double]
    L.o_set_wind.argtypes = [C.c_void_p, PD]
    L.o_set_series.argtypes = [C.c_void_p, PD, PD]
    L.o_wdflag.argtypes = [C.c_void_p, PI]
    L.o_init.argtypes = [C.c_void_p, PI]
    L.o_body_resid.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, PD]
    L.o_boundary_resid.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, PD]
    L.o_resid.argtypes = [C.c_void_p, PD]
    L.o_load.argtypes = [C.c_void_p, PD, PD]
    L.o_elem_jacobian.argtypes = [C.c_void_p, C.c_int, PD]
    L.o_linear_setup.argtypes = [C.c_void_p, C.POINTER(OOpts), PD]
    L.o_linear_solve.argtypes = [C.c_void_p, C.POINTER(OOpts), PD, PI]
    L.o_linear_solve.restype = C.c_int
    L.o_inc.argtypes = [C.c_void_p, PD]
    L.o_shift.argtypes = [C.c_void_p]
    L.o_newton.argtypes = [C.c_void_p, C.POINTER(OOpts), C.POINTER(OReport)]
    L.o_newton.restype = C.c_int
    L.o_get_residual.argtypes = [C.c_void_p, PD]
    L.o_get_matrix.argtypes = [C.c_void_p, PD, PD]
    L.o_set_matrix.argtypes = [C.c_void_p, PD, PD, PD]
    L.o_get_rhs.argtypes = [C.c_void_p, PD]
    L.o_bcgstab_only.argtypes = [C.c_void_p, C.POINTER(OOpts), PD, PI]
    L.o_bcgstab_only.restype = C.c_int
    L.o_log.argtypes = [C.c_double]
    L.o_log.restype = C.c_double
    L.o_set_log_mode.argtypes = [C.c_int]
    L.o_drag_coef.argtypes = [C.c_double, C.c_double]
    L.o_drag_coef.restype = C.c_double


def set_log_mode(mode):
    """0: platform libm log (the reference's behaviour); 1: correctly rounded."""
    lib().o_set_log_mode(mode)


def _pd(a):
    return a.ctypes.data_as(PD)


def _pi(a):
    return a.ctypes.data_as(PI)


def opts(tol_nonlin=1e-8, inc_nonlin=1e-14, max_nonlin_it=50, max_lin_it=500, max_cuts=5, prec=1, nblock=1,
         lin_rtol=1e-5, lin_atol=1e-10, hist_cap=256):
    return OOpts(tol_nonlin, inc_nonlin, max_nonlin_it, max_lin_it, max_cuts, prec, nblock, lin_rtol, lin_atol,
                 hist_cap)


class Oracle:
    """One oracle model bound to an SW2Problem (arrays are kept alive here)."""

    def __init__(self, p):
        L = lib()
        self.p = p
        self._keep = []

        def k(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self._keep.append(a)
            return a

        d = ODesc()
        d.nnodes, d.nelems2d, d.nelems1d, d.nmat = p.nnodes, p.nelems2d, p.nelems1d, p.nmat
        d.x, d.y, d.z = _pd(k(p.x, np.float64)), _pd(k(p.y, np.float64)), _pd(k(p.z, np.float64))
        d.node_bc = _pi(k(p.node_bc, np.int32))
        d.node_bc_val = _pd(k(p.node_bc_val, np.float64))
        d.e2d = _pi(k(p.e2d, np.int32))
        d.e2d_mat = _pi(k(p.e2d_mat, np.int32))
        d.e2d_ks = _pd(k(p.e2d_ks, np.float64))
        d.e2d_src_flag = _pi(k(p.e2d_src_flag, np.int32))
        d.e2d_src = _pd(k(p.e2d_src, np.float64))
        d.e1d = _pi(k(p.e1d, np.int32))
        d.e1d_adj = _pi(k(p.e1d_adj, np.int32))
        d.e1d_bc = _pi(k(p.e1d_bc, np.int32))
        d.e1d_val = _pd(k(p.e1d_val, np.float64))
        d.mat_turb = _pi(k(p.mat_turb, np.int32))
        d.mat_eev_mode = _pi(k(p.mat_eev_mode, np.int32))
        d.mat_eev_coef = _pd(k(p.mat_eev_coef, np.float64))
        d.mat_ev = _pd(k(p.mat_ev, np.float64))
        d.mat_cor_speed = _pd(k(p.mat_cor_speed, np.float64))
        d.g, d.dt, d.tau_temporal, d.tau_pg = p.g, p.dt, p.tau_temporal, p.tau_pg
        d.drying_lower, d.drying_upper, d.viscosity = p.drying_lower, p.drying_upper, p.viscosity
        d.wind = int(getattr(p, "wind", 0) or 0)
        d.density = float(getattr(p, "density", 1000.0))
        if d.wind and getattr(p, "mat_windatt", None) is not None:
            d.mat_windatt = _pd(k(p.mat_windatt, np.float64))
        self.desc = d
        self.h = L.o_create(C.byref(d))
        self.L = L
        N = p.nnodes
        self.set_state(p.head0, p.vel0, p.head0, p.vel0, p.head0, p.vel0)

    def __del__(self):
        try:
            self.L.o_destroy(self.h)
        except Exception:
            pass

    @property
    def N(self):
        return self.p.nnodes

    def set_state(self, head, vel, old_head, old_vel, older_head, older_vel):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (head, vel, old_head, old_vel, older_head, older_vel)]
        self.L.o_set_state(self.h, *[_pd(a) for a in arrs])

    def get_state(self):
        N = self.N
        out = [np.empty(N), np.empty(2 * N), np.empty(N), np.empty(2 * N), np.empty(N), np.empty(2 * N)]
        self.L.o_get_state(self.h, *[_pd(a) for a in out])
        return out

    def set_dt(self, dt):
        self.L.o_set_dt(self.h, dt)

    def set_wind(self, stress_xy):
        self.L.o_set_wind(self.h, _pd(np.ascontiguousarray(stress_xy, dtype=np.float64)))

    def set_series(self, e1d_val=None, e2d_src=None):
        a = None if e1d_val is None else np.ascontiguousarray(e1d_val, dtype=np.float64)
        b = None if e2d_src is None else np.ascontiguousarray(e2d_src, dtype=np.float64)
        self.L.o_set_series(self.h, None if a is None else _pd(a), None if b is None else _pd(b))

    def geometry(self):
        p = self.p
        dj = np.empty(p.nelems2d)
        gr = np.empty(6 * p.nelems2d)
        dj1 = np.empty(max(p.nelems1d, 1))
        nr1 = np.empty(2 * max(p.nelems1d, 1))
        self.L.o_get_geometry(self.h, _pd(dj), _pd(gr), _pd(dj1), _pd(nr1))
        return dj, gr, dj1[: p.nelems1d], nr1[: 2 * p.nelems1d]

    def shift(self):
        self.L.o_shift(self.h)

    def wdflag(self):
        out = np.empty(self.p.nelems2d, dtype=np.int32)
        self.L.o_wdflag(self.h, _pi(out))
        return out

    def init(self):
        out = np.empty(3 * self.N, dtype=np.int32)
        self.L.o_init(self.h, _pi(out))
        return out

    def body_resid(self, ie, pert=0.0, pn=0, pv=-1, ps=0):
        out = np.empty(9)
        self.L.o_body_resid(self.h, ie, pert, pn, pv, ps, _pd(out))
        return out

    def boundary_resid(self, ie, pert=0.0, pn=0, pv=-1, ps=0):
        out = np.empty(6)
        self.L.o_boundary_resid(self.h, ie, pert, pn, pv, ps, _pd(out))
        return out

    def resid(self):
        out = np.empty(3 * self.N)
        self.L.o_resid(self.h, _pd(out))
        return out

    def load(self):
        self.L.o_load(self.h, None, None)
        return self.matrix()

    def pattern(self):
        nnz = self.L.o_pattern_nnz(self.h)
        rp = np.empty(self.N + 1, dtype=np.int32)
        col = np.empty(max(nnz, 1), dtype=np.int32)
        self.L.o_get_pattern(self.h, _pi(rp), _pi(col))
        return rp, col[:nnz]

    def matrix(self):
        nnz = self.L.o_pattern_nnz(self.h)
        diag = np.empty(9 * self.N)
        offd = np.empty(9 * max(nnz, 1))
        self.L.o_get_matrix(self.h, _pd(diag), _pd(offd))
        return diag, offd[: 9 * nnz]

    def set_matrix(self, diag, offd, rhs=None):
        diag = np.ascontiguousarray(diag, dtype=np.float64)
        offd = np.ascontiguousarray(offd, dtype=np.float64)
        r = None if rhs is None else _pd(np.ascontiguousarray(rhs, dtype=np.float64))
        self.L.o_set_matrix(self.h, _pd(diag), _pd(offd), r)

    def rhs(self):
        b = np.empty(3 * self.N)
        self.L.o_get_rhs(self.h, _pd(b))
        return b

    def residual(self):
        b = np.empty(3 * self.N)
        self.L.o_get_residual(self.h, _pd(b))
        return b

    def elem_jacobian(self, ie):
        out = np.empty(81)
        self.L.o_elem_jacobian(self.h, ie, _pd(out))
        return out

    def linear_setup(self, o):
        s = np.empty(3 * self.N)
        self.L.o_linear_setup(self.h, C.byref(o), _pd(s))
        return s

    def linear_solve(self, o):
        x = np.empty(3 * self.N)
        it = C.c_int(0)
        ok = self.L.o_linear_solve(self.h, C.byref(o), _pd(x), C.byref(it))
        return ok, x, it.value

    def bcgstab_only(self, o):
        x = np.empty(3 * self.N)
        it = C.c_int(0)
        ok = self.L.o_bcgstab_only(self.h, C.byref(o), _pd(x), C.byref(it))
        return ok, x, it.value

    def inc(self, sol):
        self.L.o_inc(self.h, _pd(np.ascontiguousarray(sol, dtype=np.float64)))

    def newton(self, o):
        cap = o.hist_cap
        hm, hl, hi = np.zeros(cap), np.zeros(cap), np.zeros(cap)
        hli = np.zeros(cap, dtype=np.int32)
        r = OReport()
        r.hist_max, r.hist_l2, r.hist_inc, r.hist_lin_it = _pd(hm), _pd(hl), _pd(hi), _pi(hli)
        ok = self.L.o_newton(self.h, C.byref(o), C.byref(r))
        n = r.nhist
        return dict(ok=ok, converged=r.converged, iterations=r.iterations, lin_its=r.lin_iterations_total,
                    hist_max=hm[:n].copy(), hist_l2=hl[:n].copy(), hist_inc=hi[:n].copy(),
                    hist_lin_it=hli[:n].copy(), time_load=r.time_load, time_setup=r.time_setup,
                    time_solve=r.time_solve, time_resid=r.time_resid)
