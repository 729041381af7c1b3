"""End-to-end golden vectors: the reference's SW2 QA cases (qa/sw2/*/tri).
Each case's committed error.out.last holds the analytic-error figures that
the reference's test-case writer (src/initio/test_cases.c) printed at the
end of its run; tests/golden/qa_<case>.npz holds the deck converted to
arrays, and adhv8_kraken_amd.qa restates the writers.

  angle_nb        45-degree channel, NB VEL + NB OTW, 10 steps, NTL 1e-10
  angle_coriolis  Coriolis at 45 N with EEV turbulence, 175 steps, NIT 4 (first step cut twice)
  angle_dam       dry-bed dam break with wetting and drying, 401 steps
  angle_source    rainfall (NB SOURCE) on a closed basin, 200 steps, mass error
  angle_wind      wind setup in a closed channel (|tau| = 1 Pa, Manning 0.5 in the
                  drag law's clamped branch), 240 steps of 25 s
  wet_dry         a day-long wind event over a basin that dries on its upper 40 %,
                  then a day of relaxation (EVS 0.075, Manning 0.02, DTL 0.1), 288 steps

angle_wind's committed error.out.last is the transect error of a run to
t = 6000 s: the deck reads TC TF 1000.0, but its 40-step run ends at 7.2e-6
in the middle of a decaying seiche, while 240 steps land on the committed
3.52483555e-08 to 5e-9 relative (5975 s and 6025 s are 3% and 0.8% away),
so the run length is fixed by the golden figure itself.

wet_dry's hot start holds zero depths, which the reference reads as
-NOT_QUITE_SMALL = -FLT_EPSILON (src/structs/ssw.c:221-224,
src/include/define.h:28); with that value its three figures reproduce to
1e-10 (a reader that used -1e-6 instead ended 0.2-18 % away).

The oracle runs the reference's algorithm (skyline-LU preconditioner with
one block, OP BLK 1) and the library runs the same preconditioner on the
device (band LU of the same blocks), with the reference's BiCGSTAB stopping
rule; only rounding separates the two Newton sequences.
"""
import numpy as np
import pytest

import qa_cases

# Per case: figures compared (indices into the error.out.last numbers, and
# into our error vector) and the relative tolerances, set about an order
# above the agreement measured (oracle: 5e-12, 7e-7, 1e-10, 0, 5e-9, 8e-11;
# library on MI355X: 7e-12, 5e-8, 4e-10, 0); they reflect each deck's Newton
# tolerance (angle_coriolis stops at 1e-4 absolute and its figure is a small difference of slopes, so
# its last digits follow rounding), not a looseness of the comparison.
CASES = {
    #               ours idx   ref idx        oracle rtol  library rtol
    "angle_nb": ([0, 1, 2], [0, 2, 4], 1e-9, 1e-9),
    "angle_coriolis": ([0], [0], 1e-5, 1e-5),
    "angle_dam": ([0, 1, 2, 3], [0, 2, 4, 5], 3e-9, 3e-9),
    "angle_source": ([0], [0], 1e-6, 1e-6),
    "angle_wind": ([0], [0], 1e-7, 1e-7),
    "wet_dry": ([0, 2, 4], [0, 2, 4], 1e-9, 1e-9),
}
# Run-length overrides (see the module docstring).
TF = {"angle_wind": 6000.0}


def load(name):
    p, ctl = qa_cases.load(name)
    if name in TF:
        ctl["tf"] = TF[name]
    return p, ctl


def test_angle_wind_run_length_is_fixed_by_the_golden_figure():
    """The transect error passes the committed figure only at t = 6000 s:
    one step either side is percent-level away, and the deck's own
    TF = 1000 s is 200x away."""
    p, ctl = qa_cases.load("angle_wind")
    want = ctl["expected"][0]
    got = {}
    for tf in (1000.0, 5975.0, 6025.0):
        c = dict(ctl)
        c["tf"] = tf
        got[tf] = qa_cases.run_oracle("angle_wind", p, c)[3][0]
    assert abs(got[1000.0] / want - 1) > 10
    assert abs(got[5975.0] / want - 1) > 5e-3
    assert abs(got[6025.0] / want - 1) > 5e-3


def test_oracle_wind_setup_approaches_the_analytic_balance():
    """angle_wind: constant wind stress |tau| = 1 (rho = 1000) on a closed
    channel; the water surface tilts until g h dh/ds = |tau|/rho along the
    wind.  The mean imbalance over the transect falls with run length."""
    p, ctl = qa_cases.load("angle_wind")
    errs = []
    for tf in (1000.0, 2000.0, 4000.0):
        c = dict(ctl)
        c["tf"] = tf
        errs.append(qa_cases.run_oracle("angle_wind", p, c)[3][0])
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 3e-4 * 1.0 / ctl["density"]


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_reproduces_error_out_last(name):
    ours, ref, rtol, _ = CASES[name]
    p, ctl = load(name)
    h, v, log, err = qa_cases.run_oracle(name, p, ctl)
    np.testing.assert_allclose(err[ours], ctl["expected"][ref], rtol=rtol)


def test_fixture_comparison_is_sharp():
    """A velocity off by 1e-6 m/s at one node of angle_nb fails the check."""
    from adhv8_kraken_amd import qa
    p, ctl = qa_cases.load("angle_nb")
    h, v, log, err = qa_cases.run_oracle("angle_nb", p, ctl)
    v = v.copy()
    v[0] += 1e-6
    bad = qa.error_nb(p, h, v)
    assert not np.allclose(bad, ctl["expected"][[0, 2, 4]], rtol=CASES["angle_nb"][2], atol=0)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_library_reproduces_error_out_last(name):
    """The decks run at Courant numbers of several hundred (dt 1000 s on a
    10 m mesh for angle_nb), where 3x3 block Jacobi stalls; the library runs
    them with the deck's own preconditioner, the profile LU of OP BLK node
    blocks, on the device."""
    ours, ref, _, rtol = CASES[name]
    p, ctl = load(name)
    print(f"[qa] {name} on the device: N={p.nnodes} E={p.nelems2d}", flush=True)
    h, v, log, err = qa_cases.run_library(name, p, ctl, preconditioner=1, nblock=ctl["nblock"],
                                          lin_rtol=ctl.get("lin_rtol", 1e-5), lin_atol=1e-10)
    rel = np.abs(err[ours] - ctl["expected"][ref]) / np.abs(ctl["expected"][ref])
    print(f"[qa] {name}: steps {len(log)} attempts {sum(l[2] for l in log)} ours {err[ours]} "
          f"reference {ctl['expected'][ref]} rel {rel}", flush=True)
    np.testing.assert_allclose(err[ours], ctl["expected"][ref], rtol=rtol)
