"""The C-ABI library loads without a GPU, exports every entry point that
include/adh_sw2_hip.h declares, and its host-only entry points behave.
No compute call is made here.  CPU only."""
import ctypes as C
import os
import re
from decimal import Decimal, getcontext

import numpy as np
import pytest

from adhv8_kraken_amd import sw2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "adh_sw2_hip.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(adh_sw2_hip_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    for must in ["adh_sw2_hip_create", "adh_sw2_hip_resid", "adh_sw2_hip_load", "adh_sw2_hip_init",
                 "adh_sw2_hip_inc", "adh_sw2_hip_linear_sys_setup", "adh_sw2_hip_linear_sys_solve",
                 "adh_sw2_hip_newton", "adh_sw2_hip_batch_create", "adh_sw2_hip_batch_timestep"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    L = C.CDLL(sw2.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing


def test_library_exports_c_names_only():
    """Entry points are unmangled C symbols (extern "C")."""
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", sw2.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for n in _declared():
        assert n in syms, n


def test_default_opts_match_reference_defaults():
    """tol 1e-8 relative, BiCGSTAB max 500 its (solv_initialize.c:54), 5 line-search
    cuts (solv_initialize.c:66), stop rule 1e-10 + 1e-5 ||M^-1 b|| (solv_bcgstab.c)."""
    o = sw2.default_opts()
    assert o.max_lin_it == 500 and o.max_cuts == 5
    assert o.lin_rtol == 1e-5 and o.lin_atol == 1e-10
    assert o.precision == 0


def test_host_cr_log_is_correctly_rounded():
    L = sw2.lib()
    getcontext().prec = 50
    rng = np.random.default_rng(5)
    for x in np.exp(rng.uniform(-20, 20, 500)):
        assert sw2.cr_log(float(x)) == float(Decimal(float(x)).ln())


def test_null_arguments_are_rejected_without_touching_the_device():
    L = sw2.lib()
    assert L.adh_sw2_hip_set_opts(None, None) == -1
    assert L.adh_sw2_hip_resid(None, None) == -1
    assert L.adh_sw2_hip_pattern_nnz(None) == -1
    assert b"null" in L.adh_sw2_hip_last_error()
    L.adh_sw2_hip_destroy(None)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(sw2, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(sw2, "_lib", None)
    with pytest.raises(sw2.SW2Error):
        sw2.lib()


def test_ensemble_arguments_are_checked_before_the_device():
    """adh_sw2_hip_batch_create rejects empty ensembles, the block ILU(0)
    preconditioner, GMRES and models that disagree on the shared constants
    before any device call."""
    from adhv8_kraken_amd import mesh
    L = sw2.lib()
    keep = []
    a, b = mesh.channel_config1(), mesh.bumpy_channel(8, 4, seed=1)
    arr = (sw2.Mesh * 2)()
    arr[0] = sw2._mesh_struct(a, keep)
    arr[1] = sw2._mesh_struct(b, keep)
    h = C.c_void_p()
    assert L.adh_sw2_hip_batch_create(arr, 0, None, C.byref(h)) == -1
    o = sw2.default_opts(preconditioner=3)
    assert L.adh_sw2_hip_batch_create(arr, 2, C.byref(o), C.byref(h)) == -5
    o = sw2.default_opts(krylov=1)
    assert L.adh_sw2_hip_batch_create(arr, 2, C.byref(o), C.byref(h)) == -5
    arr[1].gravity = 9.81
    assert L.adh_sw2_hip_batch_create(arr, 2, None, C.byref(h)) == -1
    assert b"share gravity" in L.adh_sw2_hip_last_error()
    assert L.adh_sw2_hip_batch_size(None) == -1
    assert L.adh_sw2_hip_batch_timestep(None, None, None, None, 0) == -1


def test_tuning_arguments_are_checked():
    L = sw2.lib()
    t = sw2.Tuning()
    assert L.adh_sw2_hip_set_tuning(None, C.byref(t)) == -1
    assert L.adh_sw2_hip_get_tuning(None, C.byref(t)) == -1
    assert L.adh_sw2_hip_autotune(None, 3, None) == -1
