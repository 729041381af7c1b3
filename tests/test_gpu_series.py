"""Time-varying series through the ABI (adh_sw2_hip_set_series).

The reference re-averages every time series over [t, t + dt] before each
time-step attempt (sseries_setall_ts_valuesAVG, src/main/adh_func.c:542,
src/structs/sseries.c:618) and the residuals read the averages: the NB VEL
flux and the NB OTW tail elevation in fe_sw2_boundary_resid
(src/fe/sw2/fe_sw2_boundary_resid.c:212-214, 372), the rain source in
fe_sw2_body_resid (src/fe/sw2/fe_sw2_body_resid.c:728-754).  Here the
config-1 channel gets a hydrograph on its inflow edges and a rain ramp on
half of its triangles; oracle and device take the same averages each step.
"""
import numpy as np
import pytest

from gpu_helpers import bits_equal, say
from adhv8_kraken_amd import mesh
from adhv8_kraken_amd.deck import Series
from adhv8_kraken_amd.sw2 import SW2Model, default_opts, YES
from oracle.oracle import Oracle, opts as oopts, set_log_mode

pytestmark = pytest.mark.gpu

INFLOW = Series("VEL", 0, np.array([0.0, 20.0, 40.0, 60.0]), np.array([[0.05], [0.12], [0.02], [0.08]]))
RAIN = Series("SOURCE", 1, np.array([0.0, 60.0]), np.array([[0.0], [2.0e-5]]))


def _problem():
    p = mesh.channel_config1()
    E = p.nelems2d
    p.e2d_src_flag = (np.arange(E) % 2).astype(np.int32)   # rain on every other triangle
    p.e2d_src = np.zeros(E)
    return p


def _series_values(p, t, dt):
    vel = p.e1d_bc == 10
    e1 = p.e1d_val.copy()
    e1[vel] = INFLOW.average(t, t + dt)
    e2 = np.where(p.e2d_src_flag == 1, RAIN.average(t, t + dt), 0.0)
    return e1, e2


def test_time_varying_series_match_the_oracle_every_step():
    """Five 10 s steps under a hydrograph and a rain ramp: from the same
    state, the device residual equals the oracle's bit for bit at every
    step (so the averaged values reach the kernels exactly); the Newton
    steps agree to 1e-10; and the values do change the residual."""
    set_log_mode(1)
    p = _problem()
    o = Oracle(p)
    g = SW2Model(p, default_opts())
    dt = p.dt
    R0 = None
    for k in range(5):
        t = k * dt
        e1, e2 = _series_values(p, t, dt)
        say(f"[series] step {k}: inflow avg {e1[p.e1d_bc == 10][0]:.6f} rain avg {e2.max():.3e}")
        g.set_state(*o.get_state())
        o.set_series(e1, e2)
        g.set_series(e1, e2)
        o.shift(); g.shift(); o.init(); g.init()
        r_o = o.resid()
        r_g = g.resid()
        assert bits_equal(r_o, r_g), f"step {k}: residuals differ"
        if k == 1:
            # the previous step's values give a different residual: the upload matters
            e1p, e2p = _series_values(p, t - dt, dt)
            g.set_series(e1p, e2p)
            assert not bits_equal(g.resid(), r_g)
            g.set_series(e1, e2)
            g.resid()
        if R0 is None:
            R0 = np.abs(r_o).max()
        tol = 1e-8 * R0
        # linear solves to 1e-12: with the default inexact solves (1e-5 of
        # ||M^-1 b||) a late history entry is the Krylov stopping error, which
        # moves with the order of the dot-product sums; solved tightly, both
        # Newton paths agree to rounding
        g.set_opts(default_opts(tol_nonlin=tol, max_nonlin_it=30, lin_rtol=1e-12, lin_atol=0.0))
        res_o = o.newton(oopts(tol_nonlin=tol, prec=2, max_nonlin_it=30, lin_rtol=1e-12, lin_atol=0.0))
        res_g = g.newton()
        assert res_o["ok"] == 1 and res_g["verdict"] == YES
        assert res_o["iterations"] == res_g["iterations"]
        assert np.allclose(res_g["hist_max"], res_o["hist_max"], rtol=0, atol=1e-10 * R0)
        for a, b in zip(o.get_state()[:2], g.get_state()[:2]):
            assert np.abs(a - b).max() <= 1e-10 * max(1.0, np.abs(a).max())
    g.close()
