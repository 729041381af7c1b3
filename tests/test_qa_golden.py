"""End-to-end golden vectors: the reference's SW2 QA cases (qa/sw2/*/tri),
whose committed error.out.last holds the analytic-error figures the
reference printed at the end of its run (src/initio/test_cases.c).

angle_nb: 45-degree channel, NB VEL outflow and NB OTW tailwater, ten
1000 s steps, Newton to 1e-10 absolute (IP NTL) with a full skyline-LU
preconditioner (OP BLK 1).  The printed figures are mean absolute
deviations of u, v from 0.1/sqrt(2) m/s and of h from 1 m."""
import numpy as np
import pytest

import qa_cases

# The figures in error.out.last carry the run's own Newton tolerance
# (1e-10 absolute on residuals of order 1e2): agreement to 1e-9 relative
# means the same converged state, not just a similar one.
ORACLE_RTOL = 1e-9
# The library solves each Newton system with block-Jacobi BiCGSTAB instead of
# the reference's exact skyline LU, to a relative linear tolerance of 1e-12,
# so its converged state differs from the reference's by the Newton tolerance.
LIBRARY_RTOL = 1e-7


def test_angle_nb_oracle_reproduces_error_out_last():
    p, ctl = qa_cases.load("angle_nb")
    h, v, its = qa_cases.run_oracle(p, ctl)
    got = qa_cases.errors("angle_nb", p, h, v)
    want = ctl["expected"][[0, 2, 4]]  # x-vel, y-vel, head absolute errors
    np.testing.assert_allclose(got, want, rtol=ORACLE_RTOL)
    assert len(its) == 10


def test_angle_nb_fixture_detects_a_wrong_answer():
    """The comparison is sharp: a state off by 1e-6 m/s at one node fails it."""
    p, ctl = qa_cases.load("angle_nb")
    h, v, _ = qa_cases.run_oracle(p, ctl)
    v = v.copy()
    v[0] += 1e-6
    got = qa_cases.errors("angle_nb", p, h, v)
    want = ctl["expected"][[0, 2, 4]]
    assert not np.allclose(got, want, rtol=ORACLE_RTOL, atol=0)


@pytest.mark.gpu
def test_angle_nb_library_reproduces_error_out_last():
    p, ctl = qa_cases.load("angle_nb")
    print(f"[qa] angle_nb on the device: N={p.nnodes} E={p.nelems2d} steps={len(ctl['steps'])}", flush=True)
    h, v, its = qa_cases.run_library(p, ctl, lin_rtol=1e-12, lin_atol=0.0)
    got = qa_cases.errors("angle_nb", p, h, v)
    want = ctl["expected"][[0, 2, 4]]
    np.testing.assert_allclose(got, want, rtol=LIBRARY_RTOL)
