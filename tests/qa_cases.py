"""Loads the QA fixtures of tests/golden (reference decks converted by
tests/golden/make_qa_fixtures.py) and runs their time loops through the
oracle or through the HIP library."""
import os

import numpy as np

from adhv8_kraken_amd import deck, mesh

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    z = np.load(os.path.join(GOLDEN, f"qa_{name}.npz"), allow_pickle=False)
    kw = {f: z[f] for f in ["x", "y", "z", "e2d", "e2d_mat", "e2d_ks", "e2d_src_flag", "e2d_src", "e1d", "e1d_adj",
                           "e1d_bc", "e1d_val", "node_bc", "node_bc_val", "mat_turb", "mat_eev_mode",
                           "mat_eev_coef", "mat_ev", "mat_cor_speed", "head0", "vel0"]}
    kw.update({s: float(z[s]) for s in ["g", "dt", "tau_temporal", "tau_pg", "drying_lower", "drying_upper",
                                        "viscosity"]})
    p = mesh.SW2Problem(name=name, **kw)
    p.meta = dict(old_head=z["old_head"], old_vel=z["old_vel"])
    tol, itl, nit, mit, prec, nblock = z["newton"]
    ctl = dict(tol_nonlin=float(tol), inc_nonlin=float(itl), max_nonlin_it=int(nit), max_lin_it=int(mit),
               prec=int(prec), nblock=int(nblock), steps=[(float(a), float(b)) for a, b in z["steps"]],
               expected=z["expected"], test_case=str(z["test_case"]))
    return p, ctl


def run_oracle(p, ctl):
    from oracle.oracle import Oracle, opts
    o = Oracle(p)
    o.set_state(p.head0, p.vel0, p.meta["old_head"], p.meta["old_vel"], p.head0, p.vel0)
    its = []
    for (_, dt) in ctl["steps"]:
        o.set_dt(dt)
        o.shift()
        r = o.newton(opts(tol_nonlin=ctl["tol_nonlin"], inc_nonlin=ctl["inc_nonlin"],
                          max_nonlin_it=ctl["max_nonlin_it"], max_lin_it=ctl["max_lin_it"], prec=ctl["prec"],
                          nblock=ctl["nblock"]))
        assert r["ok"] == 1, ("oracle Newton failed", dt, r["hist_max"])
        its.append(r["iterations"])
    h, v, *_ = o.get_state()
    return h, v, its


def run_library(p, ctl, **opt):
    from adhv8_kraken_amd import sw2
    o = sw2.default_opts(tol_nonlin=ctl["tol_nonlin"], inc_nonlin=ctl["inc_nonlin"],
                         max_nonlin_it=ctl["max_nonlin_it"], **opt)
    m = sw2.SW2Model(p, o)
    m.set_state(p.head0, p.vel0, p.meta["old_head"], p.meta["old_vel"], p.head0, p.vel0)
    its = []
    for (_, dt) in ctl["steps"]:
        m.set_dt(dt)
        m.shift()
        r = m.newton()
        assert r["verdict"] == sw2.YES, ("library Newton failed", dt, r["hist_max"])
        its.append(r["iterations"])
    h, v, *_ = m.get_state()
    m.close()
    return h, v, its


def errors(name, p, h, v):
    if name == "angle_nb":
        return np.array(deck.testcase_error_nb(h, v, p.nnodes))
    raise KeyError(name)
