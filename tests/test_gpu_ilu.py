"""Block ILU(0) preconditioner (preconditioner 3) on the device against the
CPU statement in oracle/ilu0.py (itself pinned by tests/test_ilu0_oracle.py).

On the scaled system of a Newton step: the device's node colouring is the
greedy colouring of the pattern, and the device apply M^-1 v equals the CPU
factors' apply to 1e-12 (FP64 factors) or 1e-4 (FP32 factors); Newton with
ILU(0) meets the same tolerance as with block Jacobi, in the same number of
iterations, with fewer Krylov iterations; graph replay equals eager launches.
"""
import numpy as np
import pytest

from gpu_helpers import bits_equal, say
from adhv8_kraken_amd import mesh
from adhv8_kraken_amd.sw2 import SW2Model, default_opts, YES
from oracle.ilu0 import BlockILU0, colour_order, greedy_colors

pytestmark = pytest.mark.gpu


def _cases():
    return {"channel-40x20": mesh.channel_config1(), "floodplain-20x10": mesh.floodplain(20, 10)}


def _scaled_system(g):
    diag, offd, rhs = g.get_system()
    rp, col = g.pattern()
    N = g.N
    D = diag.reshape(N, 3, 3)
    O = {}
    for i in range(N):
        for p in range(rp[i], rp[i + 1]):
            O[(i, int(col[p]))] = offd[9 * p:9 * p + 9].reshape(3, 3)
    return D, O, rp, col


def _stepped(p, pc, precision, graphs=0):
    g = SW2Model(p, default_opts(preconditioner=pc, precision=precision, graphs=graphs))
    g.shift(); g.init(); g.resid(); g.load()
    g.linear_sys_setup()
    return g


@pytest.mark.parametrize("name", ["channel-40x20", "floodplain-20x10"])
def test_ilu0_apply_matches_the_cpu_factors(name):
    p = _cases()[name]
    say(f"ilu0 apply: {name}")
    g = _stepped(p, 3, 0)
    D, O, rp, col = _scaled_system(g)
    color, ncolor = g.ilu_colors()
    assert np.array_equal(color, greedy_colors(rp, col, g.N))
    assert all(color[int(j)] != color[i] for i in range(g.N) for j in col[rp[i]:rp[i + 1]])
    _, pos = colour_order(color)
    M = BlockILU0(D, O, pos)
    rng = np.random.default_rng(5)
    for _ in range(2):
        v = rng.uniform(-1, 1, 3 * g.N)
        y_g = g.precond_apply(v)
        y_c = M.apply(v)
        rel = np.abs(y_g - y_c).max() / np.abs(y_c).max()
        say(f"  FP64 factors: {ncolor} colours, rel.diff {rel:.2e}")
        assert rel <= 1e-12
    g.close()
    g32 = _stepped(p, 3, 1)
    v = rng.uniform(-1, 1, 3 * g32.N)
    rel = np.abs(g32.precond_apply(v) - M.apply(v)).max() / np.abs(M.apply(v)).max()
    say(f"  FP32 factors: rel.diff {rel:.2e}")
    assert rel <= 1e-4
    g32.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_newton_with_ilu0_meets_the_tolerance_in_fewer_krylov_iterations(precision):
    p = mesh.bumpy_channel(60, 30)
    res = {}
    for pc in (2, 3):
        g = SW2Model(p, default_opts(preconditioner=pc, precision=precision))
        g.shift(); g.init()
        r0 = np.abs(g.resid()).max()
        g.set_opts(default_opts(preconditioner=pc, precision=precision, tol_nonlin=1e-8 * r0, max_nonlin_it=30))
        r = g.newton()
        res[pc] = (r, g.get_state()[:2])
        g.close()
        say(f"  precision {precision} preconditioner {pc}: newton {r['iterations']} krylov {r['lin_its']}")
    (rb, sb), (ri, si) = res[2], res[3]
    assert rb["verdict"] == ri["verdict"] == YES
    assert ri["iterations"] == rb["iterations"]
    assert ri["lin_its"] < rb["lin_its"]
    for a, b in zip(sb, si):
        assert np.abs(a - b).max() <= 1e-8 * max(1.0, np.abs(a).max())


def test_ilu0_graph_replay_equals_eager():
    p = mesh.bumpy_channel(60, 30)
    out = []
    for graphs in (0, 1):
        g = SW2Model(p, default_opts(preconditioner=3, precision=1, graphs=graphs))
        g.shift(); g.init()
        r0 = np.abs(g.resid()).max()
        g.set_opts(default_opts(preconditioner=3, precision=1, graphs=graphs, tol_nonlin=1e-8 * r0,
                                max_nonlin_it=30))
        for _ in range(2):
            r = g.timestep()
            assert r["verdict"] == YES
        out.append(g.get_state()[:2])
        g.close()
    assert bits_equal(out[0][0], out[1][0]) and bits_equal(out[0][1], out[1][1])
