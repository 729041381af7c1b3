"""Writes tests/golden/qa_<case>.npz from the reference's SW2 QA decks.

Each fixture holds the deck's data as read by adhv8_kraken_amd.deck (mesh,
materials, boundary edges and values, initial state, Newton controls and
the time-step sequence) together with the figures of the case's committed
error.out.last, which the reference's test-case writer
(src/initio/test_cases.c) produced at the end of the run.

Run here, where the reference tree is mounted:
    python tests/golden/make_qa_fixtures.py [/root/reference [case ...]]
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from adhv8_kraken_amd import deck  # noqa: E402

CASES = {
    "angle_nb": "qa/sw2/angle_nb/tri/angle_nb",
    "angle_coriolis": "qa/sw2/angle_coriolis/tri/2D-coriolis_X",
    "angle_dam": "qa/sw2/angle_dam/tri/angle_dam",
    "angle_source": "qa/sw2/angle_source/tri/angle_source",
    "wet_dry": "qa/sw2/wet_dry/tri/wd",
    "angle_wind": "qa/sw2/angle_wind/tri/angle_wind",
}

FIELDS = ["x", "y", "z", "e2d", "e2d_mat", "e2d_ks", "e2d_src_flag", "e2d_src", "e1d", "e1d_adj", "e1d_bc",
          "e1d_val", "node_bc", "node_bc_val", "mat_turb", "mat_eev_mode", "mat_eev_coef", "mat_ev",
          "mat_cor_speed", "head0", "vel0"]
SCALARS = ["g", "dt", "tau_temporal", "tau_pg", "drying_lower", "drying_upper", "viscosity"]


def expected_errors(path):
    """Every floating-point figure of error.out.last, in file order."""
    txt = open(path).read()
    return np.array([float(v) for v in re.findall(r"[-+]?\d+\.\d+e[-+]\d+", txt)], dtype=np.float64), txt


def main(ref, only=None):
    for name, rel in CASES.items():
        if only and name not in only:
            continue
        stem = os.path.join(ref, rel)
        d = deck.read_deck(stem)
        p = d.problem
        out = {f: getattr(p, f) for f in FIELDS}
        out.update({s: np.float64(getattr(p, s)) for s in SCALARS})
        for k in ("old_head", "old_vel", "older_head", "older_vel"):
            out[k] = p.meta[k]
        for sid in set(d.edge_series.tolist()) - {-1}:
            if not d.series[sid].is_constant():
                raise SystemExit(f"{name}: boundary series {sid} varies in time")
        out["t0_tf"] = np.array([d.t0, d.tf])
        out["dt_series"] = np.stack([d.dt_series.times, d.dt_series.values[:, 0]])
        out["test_args"] = np.array(getattr(d, "test_args", []), dtype=np.float64)
        out["density"] = np.float64(d.density)
        out["latitude"] = np.array(p.meta["latitude"], dtype=np.float64)
        out["newton"] = np.array([d.tol_nonlin, d.inc_nonlin, d.max_nonlin_it, d.max_lin_it, d.prec, d.nblock],
                                 dtype=np.float64)
        vals, txt = expected_errors(os.path.join(os.path.dirname(stem), "error.out.last"))
        out["expected"] = vals
        out["expected_text"] = np.array(txt)
        out["test_case"] = np.array(d.test_case or "")
        out["wind"] = np.int32(p.wind)
        out["mat_windatt"] = p.mat_windatt
        winds = [sr for sr in d.series.values() if sr.kind == "WIND"]
        out["wind_series"] = (np.concatenate([winds[0].times[:, None], winds[0].values[:, :2]], axis=1)
                              if winds else np.zeros((0, 3)))
        dst = os.path.join(HERE, f"qa_{name}.npz")
        np.savez_compressed(dst, **out)
        print(dst, os.path.getsize(dst), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference", sys.argv[2:] or None)
