"""Loads the QA fixtures of tests/golden (reference decks converted by
tests/golden/make_qa_fixtures.py) and runs their time loops, with the
reference's time control (adhv8_kraken_amd.timeloop), through the oracle or
through the HIP library."""
import os

import numpy as np

from adhv8_kraken_amd import mesh, qa, timeloop

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = ["x", "y", "z", "e2d", "e2d_mat", "e2d_ks", "e2d_src_flag", "e2d_src", "e1d", "e1d_adj", "e1d_bc",
          "e1d_val", "node_bc", "node_bc_val", "mat_turb", "mat_eev_mode", "mat_eev_coef", "mat_ev",
          "mat_cor_speed", "head0", "vel0"]


def load(name):
    z = np.load(os.path.join(GOLDEN, f"qa_{name}.npz"), allow_pickle=False)
    kw = {f: z[f] for f in FIELDS}
    kw.update({s: float(z[s]) for s in ["g", "dt", "tau_temporal", "tau_pg", "drying_lower", "drying_upper",
                                        "viscosity"]})
    if "wind" in z.files:
        kw.update(wind=int(z["wind"]), mat_windatt=z["mat_windatt"], density=float(z["density"]))
    p = mesh.SW2Problem(name=name, **kw)
    p.meta = {k: z[k] for k in ("old_head", "old_vel", "older_head", "older_vel")}
    tol, itl, nit, mit, prec, nblock = z["newton"]
    ctl = dict(tol_nonlin=float(tol), inc_nonlin=float(itl), max_nonlin_it=int(nit), max_lin_it=int(mit),
               prec=int(prec), nblock=int(nblock), t0=float(z["t0_tf"][0]), tf=float(z["t0_tf"][1]),
               dt_times=z["dt_series"][0], dt_values=z["dt_series"][1], expected=z["expected"],
               test_case=str(z["test_case"]), test_args=z["test_args"], density=float(z["density"]),
               latitude=z["latitude"],
               wind_series=z["wind_series"] if "wind_series" in z.files else np.zeros((0, 3)))
    return p, ctl


def wind_at(p, ctl, t):
    """Node wind values at t (one station, weight 1): deck.Series.value_at."""
    from adhv8_kraken_amd.deck import Series
    ws = ctl["wind_series"]
    out = np.zeros(2 * p.nnodes)
    if len(ws):
        s = Series("WIND", 0, ws[:, 0], ws[:, 1:])
        if ctl["density"] > 0.0:
            out[0::2] += s.value_at(t, 0) * 1.0
            out[1::2] += s.value_at(t, 1) * 1.0
    return out


def _initial_state(p):
    m = p.meta
    return p.head0, p.vel0, m["old_head"], m["old_vel"], m["older_head"], m["older_vel"]


def make_report(name, p, ctl, djac=None):
    """Returns f(head, vel, t_prev, dt) -> the case's error figures after a step."""
    if name == "angle_nb":
        return lambda h, v, t, dt: qa.error_nb(p, h, v)
    if name == "angle_coriolis":
        n1, n2 = int(ctl["test_args"][0]) - 1, int(ctl["test_args"][1]) - 1
        return lambda h, v, t, dt: qa.error_xcoriolis(p, h, v, n1, n2, float(ctl["latitude"][0]))
    if name == "angle_dam":
        rep = qa.Dam2dError(*[float(a) for a in ctl["test_args"][:3]])
        return lambda h, v, t, dt: rep(p, h, v, t, dt)
    if name == "angle_source":
        return lambda h, v, t, dt: qa.error_water_source(p, h, ctl["density"], djac, t + dt)
    if name == "angle_wind":
        n1, n2 = int(ctl["test_args"][0]) - 1, int(ctl["test_args"][1]) - 1
        return lambda h, v, t, dt: qa.error_winds(p, h, wind_at(p, ctl, t), n1, n2, p.g, ctl["density"])
    if name == "wet_dry":
        # testcase_prep_wet_dry keeps the hot-started depth; the initial mass
        # is taken from it as the reference does at start-up
        m0 = qa.grid_mass(p, p.head0, ctl["density"], djac)
        return lambda h, v, t, dt: qa.error_wet_dry(p, h, v, p.head0, ctl["density"], djac, m0)
    raise KeyError(name)


def _run(name, p, ctl, model, get_state, djac):
    report = make_report(name, p, ctl, djac)
    last = {}

    def on_step(t, dt, info):
        if report is not None and (name == "angle_dam" or t + dt >= ctl["tf"]):
            h, v = get_state()
            last["err"] = report(h, v, t, dt)

    log = timeloop.run(model, ctl["t0"], ctl["tf"], ctl["dt_times"], ctl["dt_values"], on_step=on_step)
    h, v = get_state()
    return h, v, log, last.get("err")


def run_oracle(name, p, ctl):
    from oracle.oracle import Oracle, opts
    o = Oracle(p)
    o.set_state(*_initial_state(p))
    O = opts(tol_nonlin=ctl["tol_nonlin"], inc_nonlin=ctl["inc_nonlin"], max_nonlin_it=ctl["max_nonlin_it"],
             max_lin_it=ctl["max_lin_it"], prec=ctl["prec"], nblock=ctl["nblock"])

    class Stepper:
        def set_dt(self, dt):
            o.set_dt(dt)

        def begin_step(self, t, dt):
            if p.wind:
                o.set_wind(wind_at(p, ctl, t))

        def shift(self):
            o.shift()

        def newton_ok(self):
            r = o.newton(O)
            return r["ok"] == 1, r["iterations"]

    def state():
        h, v, *_ = o.get_state()
        return h, v

    h, v, log, err = _run(name, p, ctl, Stepper(), state, o.geometry()[0])
    return h, v, log, err


def run_library(name, p, ctl, **opt):
    from adhv8_kraken_amd import sw2
    o = sw2.default_opts(tol_nonlin=ctl["tol_nonlin"], inc_nonlin=ctl["inc_nonlin"],
                         max_nonlin_it=ctl["max_nonlin_it"], **opt)
    m = sw2.SW2Model(p, o)
    m.set_state(*_initial_state(p))

    class Stepper:
        def set_dt(self, dt):
            m.set_dt(dt)

        def begin_step(self, t, dt):
            if p.wind:
                m.set_wind(wind_at(p, ctl, t))

        def shift(self):
            m.shift()

        def newton_ok(self):
            r = m.newton()
            return r["verdict"] == sw2.YES, r["iterations"]

    def state():
        h, v, *_ = m.get_state()
        return h, v

    try:
        h, v, log, err = _run(name, p, ctl, Stepper(), state, m.geometry()[0])
    finally:
        m.close()
    return h, v, log, err
