"""The tiled band LU (the profile-LU preconditioner for node blocks of any
size, up to OP BLK 1, the reference's default: one LU of the whole matrix)
on the device against the oracle's skyline LU (solv_blk_*.c restated).

Both factor the same Jacobi-scaled system in the reference's node order
without pivoting and with the same SMALL guard, so an OP BLK 1 solve is
exact on both sides (BiCGSTAB stops after its prologue) and the solutions
agree to rounding; the Newton sequences then agree in iteration counts and
to rounding in their histories.
"""
import numpy as np
import pytest

import qa_cases
from gpu_helpers import pair, say, stepped_problem
from adhv8_kraken_amd import mesh
from adhv8_kraken_amd.sw2 import SW2Model, default_opts, YES
from oracle.oracle import Oracle, opts as oopts, set_log_mode

pytestmark = pytest.mark.gpu

CASES = {"channel-40x20": mesh.channel_config1(), "floodplain-20x10": mesh.floodplain(20, 10)}


@pytest.fixture(scope="module")
def stepped():
    return {k: stepped_problem(p) for k, p in CASES.items()}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("nblock", [1, 3])
def test_tiled_lu_linear_solve_against_the_oracle(name, nblock, stepped):
    """BiCGSTAB with the tiled LU of `nblock` node blocks against the
    oracle's skyline LU on the same scaled system (the floodplain's systems
    have near-dry rows, whose tiny pivots take the guard): one block is an
    exact solve on both sides (solutions within 1e-9); with three blocks
    both iterate to the solve's 1e-12 tolerance, and on the ill-conditioned
    floodplain system (60-80 iterations, the counts differing with the
    preconditioner's rounding) the two solutions agree within 1e-7."""
    p = CASES[name]
    say(f"tiled LU linear solve: {name} nblock={nblock}")
    o, g = pair(p, stepped[name], default_opts(preconditioner=1, nblock=nblock, lu_tiled=1, lin_rtol=1e-12,
                                               lin_atol=0.0))
    o.shift(); g.shift(); o.init(); g.init(); o.resid(); g.resid(); o.load(); g.load()
    oo = oopts(prec=1, nblock=nblock, lin_rtol=1e-12, lin_atol=0.0)
    o.linear_setup(oo)
    g.linear_sys_setup()
    ok_o, x_o, it_o = o.linear_solve(oo)
    vd, x_g, it_g = g.linear_sys_solve()
    rel = np.abs(x_g - x_o).max() / np.abs(x_o).max()
    say(f"  iterations oracle={it_o} gpu={it_g} rel.diff={rel:.3e}")
    assert vd == YES and ok_o == 1
    if nblock == 1:
        assert it_o <= 1 and it_g <= 1
    assert rel < (1e-7 if (nblock > 1 and name.startswith("flood")) else 1e-9)


def test_tiled_lu_equals_the_workgroup_lu():
    """The same blocks factored by the one-workgroup band LU and by the tiled
    LU (both exact, in the same order) give the same solution to rounding."""
    p = CASES["channel-40x20"]
    sols = []
    for tiled in (0, 1):
        g = SW2Model(p, default_opts(preconditioner=1, nblock=2, lu_tiled=tiled, lin_rtol=1e-13, lin_atol=0.0))
        g.shift(); g.init(); g.resid(); g.load()
        g.linear_sys_setup()
        vd, x, it = g.linear_sys_solve()
        assert vd == YES
        sols.append((x, it))
        g.close()
    rel = np.abs(sols[1][0] - sols[0][0]).max() / np.abs(sols[0][0]).max()
    say(f"  workgroup LU its={sols[0][1]} tiled its={sols[1][1]} rel.diff={rel:.3e}")
    assert rel < 1e-11


def test_op_blk_1_beyond_the_workgroup_lu_newton():
    """OP BLK 1 on a mesh whose single block (7,381 nodes, 22,143 DOFs) is
    beyond the one-workgroup LU, so the tiled LU is chosen by itself: one
    time step of Newton against the oracle's OP BLK 1 run: same iterations,
    every linear system solved by the LU alone on both sides (BiCGSTAB stops
    after its prologue), residual history within 1e-9 of R0 and state within
    1e-9.  The bound is the LUs' rounding, not a solve tolerance: the two
    factor in different orders, without pivoting, and the history entries
    differ by 1.1e-10 of R0 at the third entry (3.6e-7 of the entry itself;
    1e-13 of R0 at the last)."""
    set_log_mode(1)
    p = mesh.bumpy_channel(120, 60, seed=11, dt=5.0)
    say(f"OP BLK 1 newton: {p.nnodes} nodes")
    o = Oracle(p)
    g = SW2Model(p, default_opts(preconditioner=1, nblock=1))
    o.shift(); g.shift(); o.init(); g.init()
    r0 = np.abs(o.resid()).max()
    tol = 1e-8 * r0
    g.set_opts(default_opts(preconditioner=1, nblock=1, tol_nonlin=tol, max_nonlin_it=30, lin_rtol=1e-12,
                            lin_atol=0.0))
    res_o = o.newton(oopts(tol_nonlin=tol, prec=1, nblock=1, max_nonlin_it=30, lin_rtol=1e-12, lin_atol=0.0))
    res_g = g.newton()
    say(f"  oracle its={res_o['iterations']} lin={res_o.get('lin_its')} hist={res_o['hist_max']}")
    say(f"  gpu    its={res_g['iterations']} lin={res_g['lin_its']} hist={res_g['hist_max']}")
    assert res_g["verdict"] == YES and res_o["ok"] == 1
    assert res_g["iterations"] == res_o["iterations"]
    assert res_g["lin_its"] == 0
    assert np.allclose(res_g["hist_max"], res_o["hist_max"], rtol=0, atol=1e-9 * r0)
    for a, b in zip(o.get_state(), g.get_state()):
        assert np.abs(a - b).max() <= 1e-9 * max(1.0, np.abs(a).max())
    g.close()


def test_qa_deck_with_the_tiled_lu():
    """angle_nb (the reference's OP BLK 1 deck) run through the tiled LU
    reproduces its error.out.last as the one-workgroup LU does."""
    from test_qa_golden import CASES as QA
    ours, ref, _, rtol = QA["angle_nb"]
    p, ctl = qa_cases.load("angle_nb")
    h, v, log, err = qa_cases.run_library("angle_nb", p, ctl, preconditioner=1, nblock=ctl["nblock"], lu_tiled=1,
                                          lin_rtol=ctl.get("lin_rtol", 1e-5), lin_atol=1e-10)
    say(f"  angle_nb tiled LU: ours {err[ours]} reference {ctl['expected'][ref]}")
    np.testing.assert_allclose(err[ours], ctl["expected"][ref], rtol=rtol)


def test_floodplain_instance_converges_with_op_blk_1_as_the_oracle():
    """The smallest wet/dry floodplain of the 64-instance ensemble (50,176
    triangles, 40 % of the nodes dry at the start, the systems that block
    Jacobi, ILU(0) and GMRES do not solve): its first time step with OP BLK 1
    on the device against the oracle's OP BLK 1 run (committed fixture, made
    by tests/golden/make_floodplain_direct.py): the same initial residual bit
    for bit, the same Newton iteration count, the residual history within
    1e-7 of R0 and the state at 512 nodes within 1e-7 of its largest value
    (the nearly singular near-dry rows amplify the two LUs' different
    rounding; measured 3.2e-9 of R0, 2.0e-8 in depth, 1.2e-14 in velocity)."""
    import os
    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "floodplain_direct_step0.npz"))
    p = mesh.floodplain(int(f["ncx"]), int(f["ncy"]), seed=5 * 1000 + int(f["instance"]), dt=2.0)
    say(f"floodplain instance {int(f['instance'])}: {p.nnodes} nodes, {p.nelems2d} triangles, OP BLK 1")
    g = SW2Model(p, default_opts(preconditioner=1, nblock=1, max_nonlin_it=40))
    g.shift(); g.init()
    r0 = float(np.abs(g.resid()).max())
    assert r0 == float(f["r0"])
    g.set_state(p.head0, p.vel0, p.head0, p.vel0, p.head0, p.vel0)
    g.set_opts(default_opts(preconditioner=1, nblock=1, max_nonlin_it=40, tol_nonlin=1e-8 * r0))
    r = g.timestep(max_retries=0)
    h, v, *_ = g.get_state()
    g.close()
    v = np.asarray(v).reshape(-1, 2)
    idx = f["nodes"]
    dh = np.abs(r["hist_max"] - f["hist_max"]).max() / r0 if len(r["hist_max"]) == len(f["hist_max"]) else np.inf
    eh = np.abs(h[idx] - f["h"]).max() / np.abs(f["h"]).max()
    ev = max(np.abs(v[idx, 0] - f["u"]).max(), np.abs(v[idx, 1] - f["v"]).max()) / \
        max(np.abs(f["u"]).max(), np.abs(f["v"]).max())
    say(f"  oracle its={int(f['iterations'])} gpu its={r['iterations']} lin={r['lin_its']}; "
        f"history {dh:.2e} R0, depth {eh:.2e}, velocity {ev:.2e}")
    assert int(f["ok"]) == 1 and r["verdict"] == YES
    assert r["iterations"] == int(f["iterations"])
    assert dh <= 1e-7 and eh <= 1e-7 and ev <= 1e-7


def test_op_blk_1_at_1m_triangles_solves_each_newton_system():
    """OP BLK 1 on the 1000 x 500-cell channel (1.5M unknowns, 37.9 GB of band:
    more than 2^31 doubles, which the band zeroing between factorisations has
    to cover): for the first three Newton systems, the LU solution alone
    (BiCGSTAB stopped after its prologue) has a true relative residual
    |A x - b|inf / |b|inf below 1e-12 against the unscaled system."""
    import scipy.sparse as sp
    p = mesh.bumpy_channel(1000, 500, seed=2)
    N = p.nnodes
    g = SW2Model(p, default_opts(preconditioner=1, nblock=1, lin_rtol=1.0, lin_atol=0.0, max_lin_it=1))
    rp, col = g.pattern()
    nrow = np.repeat(np.arange(N), np.diff(rp))
    a = np.tile(np.repeat(np.arange(3), 3), N)
    b = np.tile(np.arange(3), 3 * N)
    rows = np.concatenate([np.repeat(3 * np.arange(N), 9) + a,
                           np.repeat(3 * nrow, 9) + np.tile(np.repeat(np.arange(3), 3), len(col))])
    cols = np.concatenate([np.repeat(3 * np.arange(N), 9) + b, np.repeat(3 * col, 9) + np.tile(np.arange(3), 3 * len(col))])
    g.shift(); g.init()
    for k in range(3):
        g.resid(); g.load()
        diag, offd, rhs = g.get_system()
        A = sp.csr_matrix((np.concatenate([diag, offd]), (rows, cols)), shape=(3 * N, 3 * N))
        g.linear_sys_setup()
        vd, x, it = g.linear_sys_solve()
        rel = np.abs(A @ x - rhs).max() / np.abs(rhs).max()
        say(f"  Newton system {k}: |Ax-b|/|b| = {rel:.2e}")
        assert it == 0 and rel < 1e-12
        g.inc(x)
    g.close()
