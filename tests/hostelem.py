"""ctypes binding of the host-compiled element code (test helper)."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "host_elem", "host_elem.cpp")
CSRC = os.path.join(os.path.dirname(HERE), "adhv8_kraken_amd", "csrc")
DEPS = [SRC, os.path.join(CSRC, "sw2_elem.h"), os.path.join(CSRC, "cr_log.h")]
LIB = os.path.join(HERE, "host_elem", "libhost_elem.so")
PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)


def build():
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-o", LIB, SRC])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if (not os.path.exists(LIB)) or os.path.getmtime(LIB) < max(os.path.getmtime(d) for d in DEPS):
            build()
        L = C.CDLL(LIB)
        L.he_create.restype = C.c_void_p
        L.he_create.argtypes = [PD] * 12 + [PI, PI, PI, PD, PD, PI, PD, PI, PD, PI, PI, PD, PD, PI, PD, PI, PI, PD,
                                            PD, PD] + [C.c_double] * 7 + [C.c_int]
        L.he_destroy.argtypes = [C.c_void_p]
        L.he_set_wind.argtypes = [C.c_void_p, PD, PD, C.c_double]
        L.he_div_rn.argtypes = [C.c_int, PD, PD, PD]
        L.he_body.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, PD]
        L.he_body_wet.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, PD]
        L.he_boundary.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, PD]
        L.he_elem_jac.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, PD]
        L.he_cr_log.argtypes = [C.c_double]
        L.he_cr_log.restype = C.c_double
        _lib = L
    return _lib


def div_rn(x, b):
    """The element code's reciprocal-corrected division, evaluated on the host."""
    x, b = _d(x), _d(b)
    q = np.empty_like(x)
    lib().he_div_rn(int(x.size), x.ctypes.data_as(PD), b.ctypes.data_as(PD), q.ctypes.data_as(PD))
    return q


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class HostElem:
    def __init__(self, p, state, wd_flag, djac, grad, djac1d, nrml1d, first_elem=None):
        head, vel, oh, ov, ooh, oov = state
        N = p.nnodes
        if first_elem is None:
            first_elem = np.full(N, -1, dtype=np.int32)
            e = p.e2d.reshape(-1, 3)
            for k in range(e.shape[0] - 1, -1, -1):
                first_elem[e[k]] = k
        self.keep = [_d(p.x), _d(p.y), _d(p.z), _d(head), _d(vel[0::2]), _d(vel[1::2]), _d(oh), _d(ov[0::2]),
                     _d(ov[1::2]), _d(ooh), _d(oov[0::2]), _d(oov[1::2]), _i(p.node_bc), _i(first_elem), _i(p.e2d),
                     _d(djac), _d(grad), _i(p.e2d_mat), _d(p.e2d_ks), _i(p.e2d_src_flag), _d(p.e2d_src),
                     _i(wd_flag), _i(p.e1d if p.nelems1d else np.zeros(2)), _d(djac1d if p.nelems1d else np.zeros(1)),
                     _d(nrml1d if p.nelems1d else np.zeros(2)), _i(p.e1d_bc if p.nelems1d else np.zeros(1)),
                     _d(p.e1d_val if p.nelems1d else np.zeros(1)), _i(p.mat_turb), _i(p.mat_eev_mode),
                     _d(p.mat_eev_coef), _d(p.mat_ev), _d(p.mat_cor_speed)]
        args = []
        for a in self.keep:
            args.append(a.ctypes.data_as(PD if a.dtype == np.float64 else PI))
        self.L = lib()
        self.h = self.L.he_create(*args, p.g, p.dt, p.tau_temporal, p.tau_pg, p.drying_lower, p.drying_upper,
                                  p.viscosity, int(p.nelems2d))
        self.p = p

    def set_wind(self, wind, mat_windatt, density):
        self._wind = [_d(wind), _d(mat_windatt)]
        self.L.he_set_wind(self.h, self._wind[0].ctypes.data_as(PD), self._wind[1].ctypes.data_as(PD), density)

    def __del__(self):
        try:
            self.L.he_destroy(self.h)
        except Exception:
            pass

    def body(self, ie, pert=0.0, pn=0, pv=-1, ps=0, dry_rule=False):
        out = np.empty(9)
        self.L.he_body(self.h, ie, pert, pn, pv, ps, int(dry_rule), out.ctypes.data_as(PD))
        return out

    def body_wet(self, ie, pert=0.0, pn=0, pv=-1, ps=0):
        out = np.empty(9)
        self.L.he_body_wet(self.h, ie, pert, pn, pv, ps, out.ctypes.data_as(PD))
        return out

    def boundary(self, ie, pert=0.0, pn=0, pv=-1, ps=0, dry_zero=False):
        out = np.empty(6)
        self.L.he_boundary(self.h, ie, pert, pn, pv, ps, int(dry_zero), out.ctypes.data_as(PD))
        return out

    def elem_jac(self, ie):
        out = np.empty(81)
        nb = self.p.node_bc
        self.L.he_elem_jac(self.h, ie, int(nb[0]), int(nb[1]), int(nb[2]), out.ctypes.data_as(PD))
        return out
