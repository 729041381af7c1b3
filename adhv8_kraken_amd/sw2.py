"""Python host mirror of the reference's SW2 Newton-path interface.

`SW2Model` wraps one device-resident model of libadh_sw2_hip.so (the C ABI in
include/adh_sw2_hip.h) with methods named after the reference functions they
replace:

    fe_sw2          -> shift()                (src/fe/sw2/fe_sw2.c:30)
    fe_sw2_init     -> init()                 (src/fe/sw2/fe_sw2_init.c:24)
    fe_sw2_resid    -> resid()                (src/fe/sw2/fe_sw2_resid.c:29)
    fe_sw2_load     -> load()                 (src/fe/sw2/fe_sw2_load.c:34)
    fe_sw2_inc      -> inc()                  (src/fe/sw2/fe_sw2_inc.c:24)
    solv_linear_sys_setup -> linear_sys_setup() (src/solver/solv_linear_sys_setup.c:15)
    solv_linear_sys_solve -> linear_sys_solve() (src/solver/solv_linear_sys_solve.c:16)
    fe_newton       -> newton()               (src/fe/fe_newton.c:214)

The library is the only compute path: if libadh_sw2_hip.so is missing or no
GPU is visible, construction raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libadh_sw2_hip.so")

YES, NO = 1, -3
PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)


class Mesh(C.Structure):
    _fields_ = [("nnodes", C.c_int), ("nelems2d", C.c_int), ("nelems1d", C.c_int), ("nmat", C.c_int),
                ("x", PD), ("y", PD), ("z", PD), ("node_bc", PI), ("node_bc_val", PD),
                ("e2d", PI), ("e2d_mat", PI), ("e2d_ks", PD), ("e2d_src_flag", PI), ("e2d_src", PD),
                ("e2d_djac", PD), ("e2d_grad", PD),
                ("e1d", PI), ("e1d_adj", PI), ("e1d_bc", PI), ("e1d_val", PD), ("e1d_djac", PD), ("e1d_nrml", PD),
                ("mat_turb", PI), ("mat_eev_mode", PI), ("mat_eev_coef", PD), ("mat_ev", PD),
                ("mat_cor_speed", PD),
                ("gravity", C.c_double), ("dt", C.c_double), ("tau_temporal", C.c_double), ("tau_pg", C.c_double),
                ("drying_lower_limit", C.c_double), ("drying_upper_limit", C.c_double),
                ("viscosity", C.c_double),
                ("wind", C.c_int), ("mat_wind_flag", PI), ("mat_windatt", PD), ("density", C.c_double)]


class Opts(C.Structure):
    _fields_ = [("tol_nonlin", C.c_double), ("inc_nonlin", C.c_double), ("max_nonlin_it", C.c_int),
                ("max_lin_it", C.c_int), ("max_cuts", C.c_int), ("lin_rtol", C.c_double), ("lin_atol", C.c_double),
                ("precision", C.c_int), ("preconditioner", C.c_int), ("check_every", C.c_int),
                ("nblock", C.c_int), ("graphs", C.c_int), ("krylov", C.c_int), ("lu_tiled", C.c_int)]


class Report(C.Structure):
    _fields_ = [("converged", C.c_int), ("iterations", C.c_int), ("lin_iterations", C.c_int),
                ("solv_fail", C.c_int), ("nhist", C.c_int), ("hist_cap", C.c_int), ("hist_max", PD),
                ("hist_l2", PD), ("hist_inc", PD), ("hist_lin_it", PI), ("ms_load", C.c_double),
                ("ms_setup", C.c_double), ("ms_solve", C.c_double), ("ms_resid", C.c_double),
                ("ms_total", C.c_double)]


class Tuning(C.Structure):
    """adh_sw2_tuning: launch shapes of the element list kernels
    (0 Jacobian wet, 1 Jacobian general, 2 residual wet, 3 residual general)."""
    _fields_ = [("block", C.c_int * 4), ("grid", C.c_int * 4), ("layout", C.c_int * 4), ("ms", C.c_double * 4)]

    def as_dict(self):
        names = ("jac_wet", "jac_gen", "res_wet", "res_gen")
        return {n: dict(block=self.block[k], grid=self.grid[k], layout=self.layout[k], ms=round(self.ms[k], 4))
                for k, n in enumerate(names)}


class SW2Error(RuntimeError):
    pass


_lib = None


def lib():
    """Loads libadh_sw2_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SW2Error(f"{LIB_PATH} is missing: build it with __graft_entry__.build() (make -C csrc)")
        L = C.CDLL(LIB_PATH)
        L.adh_sw2_hip_last_error.restype = C.c_char_p
        L.adh_sw2_hip_default_opts.argtypes = [C.POINTER(Opts)]
        L.adh_sw2_hip_create.argtypes = [C.POINTER(Mesh), C.POINTER(Opts), C.POINTER(C.c_void_p)]
        L.adh_sw2_hip_destroy.argtypes = [C.c_void_p]
        L.adh_sw2_hip_set_opts.argtypes = [C.c_void_p, C.POINTER(Opts)]
        L.adh_sw2_hip_set_dt.argtypes = [C.c_void_p, C.c_double]
        L.adh_sw2_hip_geometry.argtypes = [C.c_void_p, PD, PD, PD, PD]
        L.adh_sw2_hip_pattern_nnz.argtypes = [C.c_void_p]
        L.adh_sw2_hip_pattern.argtypes = [C.c_void_p, PI, PI]
        L.adh_sw2_hip_set_state.argtypes = [C.c_void_p] + [PD] * 6
        L.adh_sw2_hip_get_state.argtypes = [C.c_void_p] + [PD] * 6
        L.adh_sw2_hip_shift.argtypes = [C.c_void_p]
        L.adh_sw2_hip_init.argtypes = [C.c_void_p, PI, PI]
        L.adh_sw2_hip_init_noalloc.argtypes = [C.c_void_p, PI, PI]
        L.adh_sw2_hip_resid.argtypes = [C.c_void_p, PD]
        L.adh_sw2_hip_load.argtypes = [C.c_void_p, PD, PD, PD]
        L.adh_sw2_hip_elem_jacobians.argtypes = [C.c_void_p, PD, PD]
        L.adh_sw2_hip_inc.argtypes = [C.c_void_p, PD]
        L.adh_sw2_hip_linear_sys_setup.argtypes = [C.c_void_p, PD, PD]
        L.adh_sw2_hip_linear_sys_solve.argtypes = [C.c_void_p, PD, PI, PI]
        L.adh_sw2_hip_precond_apply.argtypes = [C.c_void_p, PD, PD]
        L.adh_sw2_hip_ilu_colors.argtypes = [C.c_void_p, PI, PI]
        L.adh_sw2_hip_set_system.argtypes = [C.c_void_p, PD, PD, PD]
        L.adh_sw2_hip_newton.argtypes = [C.c_void_p, C.POINTER(Report), PI]
        L.adh_sw2_hip_timestep.argtypes = [C.c_void_p, C.POINTER(Report), PI, C.c_int]
        L.adh_sw2_hip_batch_create.argtypes = [C.POINTER(Mesh), C.c_int, C.POINTER(Opts), C.POINTER(C.c_void_p)]
        L.adh_sw2_hip_batch_size.argtypes = [C.c_void_p]
        L.adh_sw2_hip_batch_set_state.argtypes = [C.c_void_p, C.c_int] + [PD] * 6
        L.adh_sw2_hip_batch_get_state.argtypes = [C.c_void_p, C.c_int, PD, PD]
        L.adh_sw2_hip_batch_set_dt.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.adh_sw2_hip_batch_timestep.argtypes = [C.c_void_p, PI, C.POINTER(Report), PI, C.c_int]
        L.adh_sw2_hip_batch_get_dt.argtypes = [C.c_void_p, C.c_int, PD]
        L.adh_sw2_hip_set_wind.argtypes = [C.c_void_p, PD]
        L.adh_sw2_hip_batch_set_wind.argtypes = [C.c_void_p, C.c_int, PD]
        L.adh_sw2_hip_set_series.argtypes = [C.c_void_p, PD, PD]
        L.adh_sw2_hip_batch_set_series.argtypes = [C.c_void_p, C.c_int, PD, PD]
        L.adh_sw2_hip_batch_set_tol.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.adh_sw2_hip_get_tuning.argtypes = [C.c_void_p, C.POINTER(Tuning)]
        L.adh_sw2_hip_set_tuning.argtypes = [C.c_void_p, C.POINTER(Tuning)]
        L.adh_sw2_hip_autotune.argtypes = [C.c_void_p, C.c_int, C.POINTER(Tuning)]
        L.adh_sw2_hip_batch_residual_norm.argtypes = [C.c_void_p, PD]
        L.adh_sw2_hip_time_kernel.argtypes = [C.c_void_p, C.c_int, C.c_int, PD]
        L.adh_sw2_hip_get_system.argtypes = [C.c_void_p, PD, PD, PD]
        L.adh_sw2_hip_probe_math.argtypes = [PD, C.c_int, PD, PD, PD]
        L.adh_sw2_hip_cr_log.argtypes = [C.c_double]
        L.adh_sw2_hip_cr_log.restype = C.c_double
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise SW2Error(f"libadh_sw2_hip error {rc}: {lib().adh_sw2_hip_last_error().decode()}")


def default_opts(**kw) -> Opts:
    o = Opts()
    lib().adh_sw2_hip_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _pd(a):
    return None if a is None else a.ctypes.data_as(PD)


def _pi(a):
    return None if a is None else a.ctypes.data_as(PI)


def _mesh_struct(p, keep):
    """adh_sw2_mesh view of a problem; the arrays it points to go into keep."""
    def k(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a

    m = Mesh()
    m.nnodes, m.nelems2d, m.nelems1d, m.nmat = p.nnodes, p.nelems2d, p.nelems1d, p.nmat
    m.x, m.y, m.z = _pd(k(p.x, np.float64)), _pd(k(p.y, np.float64)), _pd(k(p.z, np.float64))
    m.node_bc = _pi(k(p.node_bc, np.int32))
    m.node_bc_val = _pd(k(p.node_bc_val, np.float64))
    m.e2d = _pi(k(p.e2d, np.int32))
    m.e2d_mat = _pi(k(p.e2d_mat, np.int32))
    m.e2d_ks = _pd(k(p.e2d_ks, np.float64))
    m.e2d_src_flag = _pi(k(p.e2d_src_flag, np.int32))
    m.e2d_src = _pd(k(p.e2d_src, np.float64))
    if p.nelems1d:
        m.e1d = _pi(k(p.e1d, np.int32))
        m.e1d_adj = _pi(k(p.e1d_adj, np.int32))
        m.e1d_bc = _pi(k(p.e1d_bc, np.int32))
        m.e1d_val = _pd(k(p.e1d_val, np.float64))
    m.mat_turb = _pi(k(p.mat_turb, np.int32))
    m.mat_eev_mode = _pi(k(p.mat_eev_mode, np.int32))
    m.mat_eev_coef = _pd(k(p.mat_eev_coef, np.float64))
    m.mat_ev = _pd(k(p.mat_ev, np.float64))
    m.mat_cor_speed = _pd(k(p.mat_cor_speed, np.float64))
    m.gravity, m.dt, m.tau_temporal, m.tau_pg = p.g, p.dt, p.tau_temporal, p.tau_pg
    m.drying_lower_limit, m.drying_upper_limit, m.viscosity = p.drying_lower, p.drying_upper, p.viscosity
    m.wind = int(getattr(p, "wind", 0) or 0)
    m.density = float(getattr(p, "density", 1000.0))
    if m.wind and getattr(p, "mat_windatt", None) is not None:
        m.mat_windatt = _pd(k(p.mat_windatt, np.float64))
    return m


class _Tunable:
    """Launch-shape autotuning shared by single models and ensembles."""

    def autotune(self, reps: int = 3) -> dict:
        """Times the candidate launch shapes of the element list kernels on
        the current state, keeps the fastest, returns them."""
        t = Tuning()
        _check(self.L.adh_sw2_hip_autotune(self.h, reps, C.byref(t)))
        return t.as_dict()

    def get_tuning(self) -> Tuning:
        t = Tuning()
        _check(self.L.adh_sw2_hip_get_tuning(self.h, C.byref(t)))
        return t

    def set_tuning(self, t: Tuning):
        _check(self.L.adh_sw2_hip_set_tuning(self.h, C.byref(t)))


class SW2Model(_Tunable):
    """One SW2 model resident on the GPU (see module docstring)."""

    def __init__(self, problem, opts: Opts | None = None):
        L = lib()
        p = problem
        self.p = p
        self._keep = []
        m = _mesh_struct(p, self._keep)
        self.opts = opts if opts is not None else default_opts()
        h = C.c_void_p()
        _check(L.adh_sw2_hip_create(C.byref(m), C.byref(self.opts), C.byref(h)))
        self.h = h
        self.L = L
        self._keep = []  # the library copied everything it needs
        self.set_state(p.head0, p.vel0, p.head0, p.vel0, p.head0, p.vel0)

    def close(self):
        if getattr(self, "h", None):
            self.L.adh_sw2_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def N(self):
        return self.p.nnodes

    def set_opts(self, opts: Opts):
        self.opts = opts
        _check(self.L.adh_sw2_hip_set_opts(self.h, C.byref(opts)))

    def set_dt(self, dt):
        _check(self.L.adh_sw2_hip_set_dt(self.h, dt))

    def set_series(self, e1d_val=None, e2d_src=None):
        """Series values averaged over the coming step (sseries_setall_ts_valuesAVG,
        src/main/adh_func.c:542): edge values [B] and rain sources [E]; None keeps them."""
        a = None if e1d_val is None else np.ascontiguousarray(e1d_val, dtype=np.float64)
        b = None if e2d_src is None else np.ascontiguousarray(e2d_src, dtype=np.float64)
        _check(self.L.adh_sw2_hip_set_series(self.h, None if a is None else _pd(a), None if b is None else _pd(b)))

    def set_wind(self, stress_xy):
        """Node wind values (x, y per node) for the next time step."""
        _check(self.L.adh_sw2_hip_set_wind(self.h, _pd(np.ascontiguousarray(stress_xy, dtype=np.float64))))

    def set_state(self, head, vel, old_head, old_vel, older_head, older_vel):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (head, vel, old_head, old_vel, older_head, older_vel)]
        _check(self.L.adh_sw2_hip_set_state(self.h, *[_pd(x) for x in a]))

    def get_state(self):
        N = self.N
        out = [np.empty(N), np.empty(2 * N), np.empty(N), np.empty(2 * N), np.empty(N), np.empty(2 * N)]
        _check(self.L.adh_sw2_hip_get_state(self.h, *[_pd(x) for x in out]))
        return out

    def geometry(self):
        p = self.p
        dj, gr = np.empty(p.nelems2d), np.empty(6 * p.nelems2d)
        dj1, nr1 = np.empty(max(1, p.nelems1d)), np.empty(2 * max(1, p.nelems1d))
        _check(self.L.adh_sw2_hip_geometry(self.h, _pd(dj), _pd(gr), _pd(dj1), _pd(nr1)))
        return dj, gr, dj1[: p.nelems1d], nr1[: 2 * p.nelems1d]

    def pattern(self):
        nnz = self.L.adh_sw2_hip_pattern_nnz(self.h)
        rp = np.empty(self.N + 1, dtype=np.int32)
        col = np.empty(max(1, nnz), dtype=np.int32)
        _check(self.L.adh_sw2_hip_pattern(self.h, _pi(rp), _pi(col)))
        return rp, col[:nnz]

    def shift(self):
        _check(self.L.adh_sw2_hip_shift(self.h))

    def init(self):
        bc = np.empty(3 * self.N, dtype=np.int32)
        wd = np.empty(self.p.nelems2d, dtype=np.int32)
        _check(self.L.adh_sw2_hip_init(self.h, _pi(bc), _pi(wd)))
        return bc, wd

    def init_noalloc(self):
        """init through the entry point that allocates no device memory."""
        bc = np.empty(3 * self.N, dtype=np.int32)
        wd = np.empty(self.p.nelems2d, dtype=np.int32)
        _check(self.L.adh_sw2_hip_init_noalloc(self.h, _pi(bc), _pi(wd)))
        return bc, wd

    def resid(self):
        r = np.empty(3 * self.N)
        _check(self.L.adh_sw2_hip_resid(self.h, _pd(r)))
        return r

    def load(self):
        nnz = self.L.adh_sw2_hip_pattern_nnz(self.h)
        diag, offd, r = np.empty(9 * self.N), np.empty(9 * max(1, nnz)), np.empty(3 * self.N)
        _check(self.L.adh_sw2_hip_load(self.h, _pd(diag), _pd(offd), _pd(r)))
        return diag, offd[: 9 * nnz], r

    def elem_jacobians(self):
        em = np.empty(81 * self.p.nelems2d)
        em1 = np.empty(36 * max(1, self.p.nelems1d))
        _check(self.L.adh_sw2_hip_elem_jacobians(self.h, _pd(em), _pd(em1)))
        return em, em1[: 36 * self.p.nelems1d]

    def inc(self, sol=None):
        s = None if sol is None else np.ascontiguousarray(sol, dtype=np.float64)
        _check(self.L.adh_sw2_hip_inc(self.h, _pd(s)))

    def linear_sys_setup(self, x=None):
        """x: the solution array on entry (the shift the solve adds back);
        None = zero, as fe_newton sets it before every solve."""
        xs = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        s = np.empty(3 * self.N)
        _check(self.L.adh_sw2_hip_linear_sys_setup(self.h, _pd(xs), _pd(s)))
        return s

    def linear_sys_solve(self):
        x = np.empty(3 * self.N)
        it, vd = C.c_int(0), C.c_int(0)
        _check(self.L.adh_sw2_hip_linear_sys_solve(self.h, _pd(x), C.byref(it), C.byref(vd)))
        return vd.value, x, it.value

    def precond_apply(self, v):
        """M^-1 v with the preconditioner of the last set-up (profile LU or ILU(0))."""
        vv = np.ascontiguousarray(v, dtype=np.float64)
        y = np.empty(3 * self.N)
        _check(self.L.adh_sw2_hip_precond_apply(self.h, _pd(vv), _pd(y)))
        return y

    def ilu_colors(self):
        color = np.empty(self.N, dtype=np.int32)
        nc = C.c_int(0)
        _check(self.L.adh_sw2_hip_ilu_colors(self.h, _pi(color), C.byref(nc)))
        return color, nc.value

    def set_system(self, diag, offd, rhs):
        d, o, r = (np.ascontiguousarray(a, dtype=np.float64) for a in (diag, offd, rhs))
        _check(self.L.adh_sw2_hip_set_system(self.h, _pd(d), _pd(o), _pd(r)))

    def get_system(self):
        nnz = self.L.adh_sw2_hip_pattern_nnz(self.h)
        diag, offd, r = np.empty(9 * self.N), np.empty(9 * max(1, nnz)), np.empty(3 * self.N)
        _check(self.L.adh_sw2_hip_get_system(self.h, _pd(diag), _pd(offd), _pd(r)))
        return diag, offd[: 9 * nnz], r

    def time_kernel(self, which: int, reps: int = 20) -> float:
        """Average launch duration in ms (0 SpMV+M^-1, 1 FD Jacobian, 2 residual, 3 assembly,
        4 copy of the 81E-double element-matrix slab)."""
        ms = C.c_double(0.0)
        _check(self.L.adh_sw2_hip_time_kernel(self.h, which, reps, C.byref(ms)))
        return ms.value

    def _report(self, cap):
        r = Report()
        self._h = [np.zeros(cap), np.zeros(cap), np.zeros(cap), np.zeros(cap, dtype=np.int32)]
        r.hist_cap = cap
        r.hist_max, r.hist_l2, r.hist_inc, r.hist_lin_it = (_pd(self._h[0]), _pd(self._h[1]), _pd(self._h[2]),
                                                            _pi(self._h[3]))
        return r

    def _rdict(self, r, vd):
        n = min(r.nhist, r.hist_cap)
        return dict(verdict=vd, converged=r.converged, iterations=r.iterations, lin_its=r.lin_iterations,
                    hist_max=self._h[0][:n].copy(), hist_l2=self._h[1][:n].copy(), hist_inc=self._h[2][:n].copy(),
                    hist_lin_it=self._h[3][:n].copy(), ms_load=r.ms_load, ms_setup=r.ms_setup,
                    ms_solve=r.ms_solve, ms_resid=r.ms_resid, ms_total=r.ms_total)

    def newton(self, hist_cap=256):
        r = self._report(hist_cap)
        vd = C.c_int(0)
        _check(self.L.adh_sw2_hip_newton(self.h, C.byref(r), C.byref(vd)))
        return self._rdict(r, vd.value)

    def timestep(self, max_retries=0, hist_cap=256):
        r = self._report(hist_cap)
        vd = C.c_int(0)
        _check(self.L.adh_sw2_hip_timestep(self.h, C.byref(r), C.byref(vd), max_retries))
        return self._rdict(r, vd.value)


class SW2Ensemble(_Tunable):
    """Independent SW2 models advanced together on one GPU: one device mesh
    holds them side by side, so every phase is one launch over all models,
    while Newton, Krylov and time-step decisions stay per model (each model
    computes what it computes alone).  3x3 block Jacobi, or the profile LU
    of each model's node blocks (opts.nblock per model, tiled band LU)."""

    def __init__(self, problems, opts: Opts | None = None):
        L = lib()
        self.ps = list(problems)
        keep = []
        arr = (Mesh * len(self.ps))()
        for i, p in enumerate(self.ps):
            arr[i] = _mesh_struct(p, keep)
        self.opts = opts if opts is not None else default_opts()
        h = C.c_void_p()
        _check(L.adh_sw2_hip_batch_create(arr, len(self.ps), C.byref(self.opts), C.byref(h)))
        self.h = h
        self.L = L
        for i, p in enumerate(self.ps):
            self.set_state(i, p.head0, p.vel0, p.head0, p.vel0, p.head0, p.vel0)

    def __len__(self):
        return len(self.ps)

    def close(self):
        if getattr(self, "h", None):
            self.L.adh_sw2_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_opts(self, opts: Opts):
        self.opts = opts
        _check(self.L.adh_sw2_hip_set_opts(self.h, C.byref(opts)))

    def set_dt(self, i, dt):
        _check(self.L.adh_sw2_hip_batch_set_dt(self.h, i, dt))

    def set_series(self, i, e1d_val=None, e2d_src=None):
        a = None if e1d_val is None else np.ascontiguousarray(e1d_val, dtype=np.float64)
        b = None if e2d_src is None else np.ascontiguousarray(e2d_src, dtype=np.float64)
        _check(self.L.adh_sw2_hip_batch_set_series(self.h, i, None if a is None else _pd(a),
                                                   None if b is None else _pd(b)))

    def set_wind(self, i, stress_xy):
        _check(self.L.adh_sw2_hip_batch_set_wind(self.h, i, _pd(np.ascontiguousarray(stress_xy, dtype=np.float64))))

    def get_dt(self, i):
        dt = C.c_double(0.0)
        _check(self.L.adh_sw2_hip_batch_get_dt(self.h, i, C.byref(dt)))
        return dt.value

    def set_tol(self, i, tol_nonlin):
        _check(self.L.adh_sw2_hip_batch_set_tol(self.h, i, tol_nonlin))

    def residual_norm(self):
        """max-norm of each model's opening Newton residual (applies init)."""
        out = np.empty(len(self.ps))
        _check(self.L.adh_sw2_hip_batch_residual_norm(self.h, _pd(out)))
        return out

    def set_state(self, i, head, vel, old_head, old_vel, older_head, older_vel):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (head, vel, old_head, old_vel, older_head, older_vel)]
        _check(self.L.adh_sw2_hip_batch_set_state(self.h, i, *[_pd(x) for x in a]))

    def get_state(self, i):
        N = self.ps[i].nnodes
        hd, vel = np.empty(N), np.empty(2 * N)
        _check(self.L.adh_sw2_hip_batch_get_state(self.h, i, _pd(hd), _pd(vel)))
        return hd, vel

    def timestep(self, max_retries=0, hist_cap=64, which=None):
        """One time step of every model (or of those with which[i] true)."""
        n = len(self.ps)
        sel = None if which is None else np.ascontiguousarray(np.asarray(which, dtype=np.int32))
        reps = (Report * n)()
        hists = []
        for i in range(n):
            hh = [np.zeros(hist_cap), np.zeros(hist_cap), np.zeros(hist_cap), np.zeros(hist_cap, dtype=np.int32)]
            hists.append(hh)
            reps[i].hist_cap = hist_cap
            reps[i].hist_max, reps[i].hist_l2, reps[i].hist_inc, reps[i].hist_lin_it = (
                _pd(hh[0]), _pd(hh[1]), _pd(hh[2]), _pi(hh[3]))
        vd = np.zeros(n, dtype=np.int32)
        _check(self.L.adh_sw2_hip_batch_timestep(self.h, _pi(sel), reps, _pi(vd), max_retries))
        out = []
        for i in range(n):
            r = reps[i]
            k = min(r.nhist, hist_cap)
            out.append(dict(verdict=int(vd[i]), converged=r.converged, iterations=r.iterations,
                            lin_its=r.lin_iterations, hist_max=hists[i][0][:k].copy(),
                            hist_l2=hists[i][1][:k].copy(), hist_inc=hists[i][2][:k].copy(),
                            hist_lin_it=hists[i][3][:k].copy()))
        return out


def profile_lu_bytes(p, nblock: int = 1) -> int:
    """Device bytes the tiled band LU takes for problem p cut into `nblock`
    node blocks by the reference's serial partition (node_order.c:64-104:
    N / nblock consecutive nodes per block, the last taking the rest).  A
    block of n nodes has m = 3 n DOFs and pattern half bandwidth w = 3 d + 2
    (d: the largest node-number distance of two nodes of one element inside
    the block); it stores the band, m (2 (w + 64) + 1) doubles, the inverses
    of its 64 x 64 diagonal factor tiles (2 x 4096 doubles per 64 DOFs), two
    band-ordered vectors and one panel copy (8,256 doubles)."""
    N = int(p.nnodes)
    nb = min(max(1, int(nblock)), N)
    inc = max(1, N // nb)
    blk = np.minimum(np.arange(N) // inc, nb - 1)
    maxd = np.zeros(nb, dtype=np.int64)
    e2 = np.asarray(p.e2d, dtype=np.int64).reshape(-1, 3)
    e1 = np.asarray(p.e1d, dtype=np.int64).reshape(-1, 2)
    for a, b in [(e2[:, 0], e2[:, 1]), (e2[:, 1], e2[:, 2]), (e2[:, 0], e2[:, 2]), (e1[:, 0], e1[:, 1])]:
        same = blk[a] == blk[b]
        if np.any(same):
            np.maximum.at(maxd, blk[a[same]], np.abs(a[same] - b[same]))
    n = np.bincount(blk, minlength=nb).astype(np.int64)
    m = 3 * n
    w = 3 * maxd + 2
    doubles = m * (2 * (w + 64) + 1) + 2 * 4096 * ((m + 63) // 64) + 2 * m + 8256
    return int(8 * doubles.sum())


def lu_groups(problems, nblock: int = 1, budget: float = 200e9) -> list:
    """Indices of `problems` in groups for ensembles with the profile LU,
    whose bands must fit in device memory together: by increasing storage
    (instances of like size share a group, so a group's launches, as many
    as its largest band has panels, serve all its members), each group's
    profile_lu_bytes within `budget` (an instance above it alone)."""
    sizes = [profile_lu_bytes(p, nblock) for p in problems]
    groups, cur, tot = [], [], 0
    for i in sorted(range(len(problems)), key=lambda k: sizes[k]):
        if cur and tot + sizes[i] > budget:
            groups.append(cur)
            cur, tot = [], 0
        cur.append(i)
        tot += sizes[i]
    if cur:
        groups.append(cur)
    return groups


def cr_log(x: float) -> float:
    return lib().adh_sw2_hip_cr_log(x)
