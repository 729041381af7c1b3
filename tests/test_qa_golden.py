"""End-to-end golden vectors: the reference's SW2 QA cases (qa/sw2/*/tri).
Each case's committed error.out.last holds the analytic-error figures that
the reference's test-case writer (src/initio/test_cases.c) printed at the
end of its run; tests/golden/qa_<case>.npz holds the deck converted to
arrays, and adhv8_kraken_amd.qa restates the writers.

  angle_nb        45-degree channel, NB VEL + NB OTW, 10 steps, NTL 1e-10
  angle_coriolis  Coriolis at 45 N with EEV turbulence, 175 steps, NIT 4 (first step cut twice)
  angle_dam       dry-bed dam break with wetting and drying, 401 steps
  angle_source    rainfall (NB SOURCE) on a closed basin, 200 steps, mass error

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
# above the agreement measured (oracle: 5e-12, 7e-7, 1e-10, 0; library on
# MI355X: 7e-12, 5e-8, 4e-10, 0); they reflect each deck's Newton tolerance (angle_coriolis
# stops at 1e-4 absolute and its figure is a small difference of slopes, so
# its last digits follow rounding), not a looseness of the comparison.
CASES = {
    #               ours idx   ref idx        oracle rtol  library rtol
    "angle_nb": ([0, 1, 2], [0, 2, 4], 1e-9, 1e-9),
    "angle_coriolis": ([0], [0], 1e-5, 1e-5),
    "angle_dam": ([0, 1, 2, 3], [0, 2, 4, 5], 3e-9, 3e-9),
    "angle_source": ([0], [0], 1e-6, 1e-6),
}


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_reproduces_error_out_last(name):
    ours, ref, rtol, _ = CASES[name]
    p, ctl = qa_cases.load(name)
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
    p, ctl = qa_cases.load(name)
    print(f"[qa] {name} on the device: N={p.nnodes} E={p.nelems2d}", flush=True)
    h, v, log, err = qa_cases.run_library(name, p, ctl, preconditioner=1, nblock=ctl["nblock"],
                                          lin_rtol=ctl.get("lin_rtol", 1e-5), lin_atol=1e-10)
    rel = np.abs(err[ours] - ctl["expected"][ref]) / np.abs(ctl["expected"][ref])
    print(f"[qa] {name}: steps {len(log)} attempts {sum(l[2] for l in log)} ours {err[ours]} "
          f"reference {ctl['expected'][ref]} rel {rel}", flush=True)
    np.testing.assert_allclose(err[ours], ctl["expected"][ref], rtol=rtol)
