"""GPU parity: libadh_sw2_hip.so on the device against the CPU oracle.

Bar (integer/index work and the element arithmetic): bit-exact.  The oracle
runs with its correctly rounded log (the device log is correctly rounded;
glibc's misrounds ~7e-4 of inputs -- see test_oracle_pins.py).  Krylov
iterates differ from the oracle's by the order of dot-product sums only
(tree vs sequential), so solutions and Newton histories are compared with a
tolerance written in each test.
"""
import numpy as np
import pytest

from gpu_helpers import bits_equal, nmismatch, pair, say, stepped_problem
from adhv8_kraken_amd import mesh
from adhv8_kraken_amd.sw2 import SW2Model, default_opts, YES
from oracle.oracle import Oracle, opts as oopts, set_log_mode

pytestmark = pytest.mark.gpu


def _cases():
    p1 = mesh.channel_config1()
    f = mesh.floodplain(20, 10)
    b = mesh.bumpy_channel(60, 30, seed=11, dt=5.0)
    return {"channel-40x20": p1, "floodplain-20x10": f, "bumpy-60x30": b}


CASES = _cases()


@pytest.fixture(scope="module")
def stepped():
    out = {}
    for k, p in CASES.items():
        out[k] = stepped_problem(p)
    return out


def test_device_rounding_probe():
    """Device cr_log equals the host's and binary128's; device sqrt and 1/x
    equal the host's (IEEE correctly rounded)."""
    import ctypes as C
    from adhv8_kraken_amd.sw2 import lib
    say("probe: device log/sqrt/div on 2^20 inputs")
    rng = np.random.default_rng(3)
    n = 1 << 20
    bits = (rng.integers(0, 1 << 52, n, dtype=np.int64) | ((rng.integers(900, 1150, n, dtype=np.int64)) << 52))
    x = bits.view(np.float64).copy()
    x[: 1000] = 1.0 + 29.7 * rng.uniform(0.01, 200.0, 1000)  # the friction law's log arguments
    lg, sq, dv = np.empty(n), np.empty(n), np.empty(n)
    PD = C.POINTER(C.c_double)
    L = lib()
    L.adh_sw2_hip_probe_math.argtypes = [PD, C.c_int, PD, PD, PD]
    assert L.adh_sw2_hip_probe_math(x.ctypes.data_as(PD), n, lg.ctypes.data_as(PD), sq.ctypes.data_as(PD),
                                    dv.ctypes.data_as(PD)) == 0
    host_log = np.array([L.adh_sw2_hip_cr_log(float(v)) for v in x[:20000]])
    assert bits_equal(lg[:20000], host_log)
    assert bits_equal(sq, np.sqrt(x))
    assert bits_equal(dv, 1.0 / x)


@pytest.mark.parametrize("name", list(CASES))
def test_geometry_and_pattern(name):
    p = CASES[name]
    say(f"geometry/pattern: {name}")
    o, g = pair(p)
    for a, b in zip(o.geometry(), g.geometry()):
        assert bits_equal(a, b)
    o.shift(); o.init(); o.resid(); o.load()
    rp_o, col_o = o.pattern()
    rp_g, col_g = g.pattern()
    assert np.array_equal(rp_o, rp_g) and np.array_equal(col_o, col_g)


@pytest.mark.parametrize("name", list(CASES))
def test_resid_load_setup_bitexact(name, stepped):
    """init -> resid -> load -> linear_sys_setup on identical states: residual,
    wet/dry flags, bc mask, post-residual velocities, assembled 3x3 blocks,
    element Jacobians, scale vector and scaled system are bit-identical."""
    p = CASES[name]
    st = stepped[name]
    say(f"resid/load/setup: {name}")
    o, g = pair(p, st)
    o.shift(); g.shift()
    bc_o = o.init()
    bc_g, wd_g = g.init()
    assert np.array_equal(bc_o, bc_g)
    assert np.array_equal(o.wdflag(), wd_g)
    r_o = o.resid()
    r_g = g.resid()
    assert nmismatch(r_o, r_g) == 0
    for a, b in zip(o.get_state(), g.get_state()):
        assert bits_equal(a, b)
    d_o, f_o = o.load()
    d_g, f_g, rr_g = g.load()
    assert nmismatch(d_o, d_g) == 0
    assert nmismatch(f_o, f_g) == 0
    assert bits_equal(o.residual(), rr_g)
    em, _ = g.elem_jacobians()
    for ie in range(0, p.nelems2d, max(1, p.nelems2d // 200)):
        assert bits_equal(o.elem_jacobian(ie), em[81 * ie: 81 * ie + 81])
    s_o = o.linear_setup(oopts(prec=2))
    s_g = g.linear_sys_setup()
    assert bits_equal(s_o, s_g)


def test_dry_node_velocity_rule():
    """Velocities at nodes with old_head <= 0 are seen only by the lowest
    element touching the node, then zeroed (fe_sw2_resid.c:85-93)."""
    p = mesh.floodplain(20, 10)
    rng = np.random.default_rng(1)
    N = p.nnodes
    h0 = p.head0.copy()
    vel = rng.uniform(-0.2, 0.2, 2 * N)
    st = (h0, vel, h0, vel, h0, vel)
    say("dry-node velocity rule: floodplain 20x10 with random velocities")
    o, g = pair(p, st)
    o.init(); g.init()
    assert nmismatch(o.resid(), g.resid()) == 0
    for a, b in zip(o.get_state(), g.get_state()):
        assert bits_equal(a, b)


@pytest.mark.parametrize("name", ["channel-40x20", "bumpy-60x30"])
@pytest.mark.parametrize("precision", [0, 1, 2])
def test_linear_solve(name, precision, stepped):
    """BiCGSTAB with 3x3 block-Jacobi on the same scaled system as the oracle.
    FP64: iteration count within 1 and solution within 1e-9 relative (dot
    sums are reordered).  FP32: solution within 1e-4 relative of FP64."""
    p = CASES[name]
    say(f"linear solve: {name} precision={precision}")
    o, g = pair(p, stepped[name], default_opts(precision=precision))
    o.shift(); g.shift(); o.init(); g.init(); o.resid(); g.resid(); o.load(); g.load()
    oo = oopts(prec=2)
    o.linear_setup(oo)
    g.linear_sys_setup()
    ok_o, x_o, it_o = o.linear_solve(oo)
    vd, x_g, it_g = g.linear_sys_solve()
    assert vd == YES and ok_o == 1
    rel = np.abs(x_g - x_o).max() / np.abs(x_o).max()
    say(f"  iterations oracle={it_o} gpu={it_g} rel.diff={rel:.3e}")
    if precision == 0:
        assert abs(it_o - it_g) <= 1
        assert rel < 1e-9
    else:
        assert rel < 1e-4


@pytest.mark.parametrize("precision", [0, 1])
def test_setup_twice_is_refused_and_load_rearms_it(precision, stepped):
    """The set-up scales the system in place; a second set-up of the same
    system is refused with an error (the refined forms keep the scaled
    blocks in FP32 only), and a new load gives it a fresh system that solves
    exactly as the first did: the set-up's x (None = zero, as fe_newton
    passes it) replaces the previous solution as the solve's shift."""
    from adhv8_kraken_amd.sw2 import SW2Error
    name = "bumpy-60x30"
    p = CASES[name]
    say(f"set-up guard: {name} precision={precision}")
    g = SW2Model(p, default_opts(precision=precision))
    g.set_state(*stepped[name])
    g.shift(); g.init(); g.resid(); g.load()
    s1 = g.linear_sys_setup()
    with pytest.raises(SW2Error, match="already scaled"):
        g.linear_sys_setup()
    vd1, x1, it1 = g.linear_sys_solve()
    g.init(); g.resid(); g.load()
    s2 = g.linear_sys_setup()
    vd2, x2, it2 = g.linear_sys_solve()
    # a caller's x is the shift the solve adds back (solv_return_old.c:23-25)
    g.init(); g.resid(); g.load()
    g.linear_sys_setup(x=x1)
    vd3, x3, it3 = g.linear_sys_solve()
    g.close()
    assert vd1 == vd2 == vd3 == YES and it1 == it2 == it3
    assert bits_equal(s1, s2) and bits_equal(x1, x2)
    assert np.abs(x3 - 2.0 * x1).max() <= 1e-14 * np.abs(x1).max()


@pytest.mark.parametrize("nblock", [1, 3])
def test_linear_solve_profile_lu(nblock, stepped):
    """BiCGSTAB with the profile-LU preconditioner of `nblock` node blocks
    (OP PRE 1 / OP BLK) against the oracle's skyline LU on the same scaled
    system: with one block the LU is exact, so both stop after the prologue
    and agree to rounding; with three blocks both iterate and agree within
    the solve's own tolerance."""
    name = "channel-40x20"
    p = CASES[name]
    say(f"profile-LU linear solve: {name} nblock={nblock}")
    o, g = pair(p, stepped[name], default_opts(preconditioner=1, nblock=nblock, lin_rtol=1e-12, lin_atol=0.0))
    o.shift(); g.shift(); o.init(); g.init(); o.resid(); g.resid(); o.load(); g.load()
    oo = oopts(prec=1, nblock=nblock, lin_rtol=1e-12, lin_atol=0.0)
    o.linear_setup(oo)
    g.linear_sys_setup()
    ok_o, x_o, it_o = o.linear_solve(oo)
    vd, x_g, it_g = g.linear_sys_solve()
    assert vd == YES and ok_o == 1
    rel = np.abs(x_g - x_o).max() / np.abs(x_o).max()
    say(f"  iterations oracle={it_o} gpu={it_g} rel.diff={rel:.3e}")
    if nblock == 1:
        assert it_o <= 1 and it_g <= 1
    assert rel < 1e-9


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_krylov_graph_replays_match_eager(precision, stepped):
    """The Krylov iterations replay captured HIP graphs: several solves in a
    row (each replaying graphs of different chunk lengths, with host
    synchronisation and eager kernels between replays) give bit-identical
    solutions and iteration counts to eager launches of the same kernels."""
    name = "bumpy-60x30"
    p = CASES[name]
    say(f"graph replays: {name} precision={precision}")
    res = {}
    for graphs in (0, 1):
        g = SW2Model(p, default_opts(precision=precision, graphs=graphs))
        g.set_state(*stepped[name])
        g.shift(); g.init(); g.resid(); g.load()
        diag, offd, rhs = g.get_system()
        out = []
        for rep in range(4):
            g.set_system(diag, offd, rhs)
            g.linear_sys_setup()
            vd, x, it = g.linear_sys_solve()
            assert vd == YES
            out.append((x, it))
        g.close()
        res[graphs] = out
    for (xe, ie), (xg, ig) in zip(res[0], res[1]):
        assert ie == ig
        assert bits_equal(xe, xg)


@pytest.mark.parametrize("name", ["channel-40x20", "bumpy-60x30"])
def test_newton_history(name):
    """One time step of Newton: the device (3x3 block-Jacobi BiCGSTAB)
    against the oracle in the same mode and in the reference's own mode,
    OP PRE 1 with OP BLK = nnodes (one node per skyline-LU block, the same
    preconditioner up to rounding): same iteration count, residual history
    within 1e-10 of R0 (H7: the order of dot/l2 sums and of the 3x3 inverse
    are the only differences; measured 7e-12 of R0 against the same mode)."""
    p = CASES[name]
    say(f"newton: {name}")
    o, g = pair(p)
    o1 = Oracle(p)
    for m in (o, g, o1):
        m.shift()
        m.init()
    r0 = np.abs(o.resid()).max()
    o1.resid()
    tol = 1e-8 * r0
    go = default_opts(tol_nonlin=tol, max_nonlin_it=30)
    g.set_opts(go)
    res_o = o.newton(oopts(tol_nonlin=tol, prec=2, max_nonlin_it=30))
    res_1 = o1.newton(oopts(tol_nonlin=tol, prec=1, nblock=p.nnodes, max_nonlin_it=30))
    res_g = g.newton()
    say(f"  oracle its={res_o['iterations']} hist={res_o['hist_max']}")
    say(f"  blk=N  its={res_1['iterations']} hist={res_1['hist_max']}")
    say(f"  gpu    its={res_g['iterations']} hist={res_g['hist_max']}")
    say(f"  max |dhist|/R0: same mode {np.abs(res_g['hist_max'] - res_o['hist_max']).max() / r0:.2e}, "
        f"prec 1 {np.abs(res_g['hist_max'] - res_1['hist_max']).max() / r0:.2e}")
    assert res_g["verdict"] == YES and res_o["ok"] == 1 and res_1["ok"] == 1
    assert res_g["iterations"] == res_o["iterations"] == res_1["iterations"]
    assert np.allclose(res_g["hist_max"], res_o["hist_max"], rtol=0, atol=1e-10 * r0)
    assert np.allclose(res_g["hist_max"], res_1["hist_max"], rtol=0, atol=1e-10 * r0)
    h_o = o.get_state()
    h_g = g.get_state()
    for a, b in zip(h_o, h_g):
        assert np.abs(a - b).max() <= 1e-10 * max(1.0, np.abs(a).max())


def test_large_mesh_jacobian_is_directional_derivative():
    """At 2.2M triangles (8,594 blocks of 256 triangles, so each of the 1,024
    threads of the list scan owns more than 8 block counts: its chunked
    path; 17,221 assembly workgroups) the assembled Jacobian applied to a
    random direction equals -dR along it, by a central difference of the
    device residual (fe_sw2_load.c:34-269 against fe_sw2_resid.c:29-216).
    The element blocks are first filled at a far-away state, so a triangle
    missing from the element lists, or a slot the assembly skipped, keeps
    blocks that are wrong by O(1) in its rows.  (Meshes below 2.1M
    triangles run the scan's register path: every other test.)"""
    import scipy.sparse
    p = mesh.bumpy_channel(1100, 1000, seed=5)
    N = p.nnodes
    say(f"large-mesh Jacobian: {p.nelems2d} triangles, {N} nodes")
    g = SW2Model(p)
    g.shift(); g.init()
    r0 = np.abs(g.resid()).max()
    g.set_opts(default_opts(tol_nonlin=1e-3 * r0, max_nonlin_it=3))
    g.newton()                                  # a state with flow
    st = g.get_state()
    g.set_state(1.5 * st[0], 3.0 * st[1], *st[2:])
    g.resid(); g.load()                         # element blocks of another state
    g.set_state(*st)
    g.resid(); g.load()
    diag, offd, _ = g.get_system()
    rp, col = g.pattern()
    bi = np.concatenate([np.arange(N), np.repeat(np.arange(N), np.diff(rp))])
    bj = np.concatenate([np.arange(N), col])
    rr = np.broadcast_to(3 * bi[:, None, None] + np.arange(3)[None, :, None], (len(bi), 3, 3))
    cc = np.broadcast_to(3 * bj[:, None, None] + np.arange(3)[None, None, :], (len(bi), 3, 3))
    A = scipy.sparse.csr_matrix((np.concatenate([diag, offd]), (rr.ravel(), cc.ravel())), shape=(3 * N, 3 * N))
    rng = np.random.default_rng(1)
    d = rng.standard_normal(3 * N)
    d[2::3] *= 1e-2
    t = 1e-6

    def R_at(s):
        h = st[0] + s * d[2::3]
        vel = st[1].copy()
        vel[0::2] += s * d[0::3]
        vel[1::2] += s * d[1::3]
        g.set_state(h, vel, *st[2:])
        return g.resid()

    fd = -(R_at(t) - R_at(-t)) / (2 * t)
    Jd = A @ d
    g.close()
    worst = np.abs(Jd - fd).max() / np.abs(fd).max()
    l2 = np.linalg.norm(Jd - fd) / np.linalg.norm(fd)
    say(f"  max row error / max |dR| = {worst:.3e}, l2 {l2:.3e}")
    assert l2 < 1e-6
    assert worst < 1e-4


def _hub_basin(nring=20, seed=3):
    """A 110 m square basin whose centre node has `nring` neighbours (a ring
    of nodes around it, Delaunay-triangulated with a 10 m grid outside),
    flume-style inflow and tailwater strings, a wavy bed and random
    velocities: rows longer than the assembly's 15 off-diagonal waves and
    than the unrolled SpMV widths."""
    import scipy.spatial
    rng = np.random.default_rng(seed)
    gx, gy = np.meshgrid(np.arange(12) * 10.0, np.arange(12) * 10.0, indexing="ij")
    pts = np.stack([gx.ravel(), gy.ravel()], 1)
    c = np.array([55.0, 55.0])
    pts = pts[np.linalg.norm(pts - c, axis=1) > 25.0]
    ang = 2 * np.pi * (np.arange(nring) + 0.5) / nring
    ring = c + 12.0 * np.stack([np.cos(ang), np.sin(ang)], 1)
    pts = np.concatenate([c[None], ring, pts])
    tri = scipy.spatial.Delaunay(pts).simplices.astype(np.int64)
    a, b, cc = pts[tri[:, 0]], pts[tri[:, 1]], pts[tri[:, 2]]
    cross = (b[:, 0] - a[:, 0]) * (cc[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (cc[:, 0] - a[:, 0])
    tri[cross < 0] = tri[cross < 0][:, [0, 2, 1]]
    N, E = len(pts), len(tri)
    x, y = pts[:, 0].copy(), pts[:, 1].copy()
    z = -(1e-3 * x) + 0.05 * np.sin(x / 17.0) * np.cos(y / 13.0)
    # flume-style strings: NB VEL inflow on the west edge, NB OTW tailwater on
    # the east edge, each edge from north to south, with its triangle
    def edges(side_x):
        nodes = np.where(np.isclose(x, side_x))[0]
        nodes = nodes[np.argsort(-y[nodes])]
        e = np.stack([nodes[:-1], nodes[1:]], 1)
        adj = [int(np.where((tri == a).any(1) & (tri == b).any(1))[0][0]) for a, b in e]
        return e, np.array(adj)
    ew, aw = edges(0.0)
    ee, ae = edges(110.0)
    tail = float(z[np.isclose(x, 110.0)].max()) + 1.0
    e1d = np.concatenate([ew, ee]).astype(np.int32).ravel()
    e1d_adj = np.concatenate([aw, ae]).astype(np.int32)
    e1d_bc = np.concatenate([np.full(len(ew), mesh.BCT_VEL_NEU), np.full(len(ee), mesh.BCT_PRS_NEU)]).astype(np.int32)
    e1d_val = np.concatenate([np.full(len(ew), 0.05), np.full(len(ee), tail)])
    p = mesh.SW2Problem(
        name="hub-basin", x=x, y=y, z=z, e2d=np.ascontiguousarray(tri.ravel().astype(np.int32)),
        e2d_mat=np.zeros(E, dtype=np.int32), e2d_ks=np.full(E, mesh.manning_to_ks(0.03)),
        e2d_src_flag=np.zeros(E, dtype=np.int32), e2d_src=np.zeros(E),
        e1d=np.ascontiguousarray(e1d), e1d_adj=np.ascontiguousarray(e1d_adj), e1d_bc=np.ascontiguousarray(e1d_bc),
        e1d_val=np.ascontiguousarray(e1d_val), node_bc=np.zeros(N, dtype=np.int32), node_bc_val=np.zeros(3 * N),
        mat_turb=np.array([mesh.TURB_EVS], dtype=np.int32), mat_eev_mode=np.array([0], dtype=np.int32),
        mat_eev_coef=np.array([0.0]), mat_ev=np.zeros(3), mat_cor_speed=np.array([0.0]), dt=5.0,
        head0=np.ascontiguousarray(tail - z), vel0=np.zeros(2 * N))
    vel = np.zeros(2 * N)
    vel[0::2] = 0.1 + 0.02 * rng.standard_normal(N)
    vel[1::2] = 0.02 * rng.standard_normal(N)
    return p, vel


def test_high_valence_mesh():
    """Residual, element Jacobians, assembled blocks and scaled system
    bit-exact, and Newton within the FP64 tolerances, on a mesh with a node
    of 20 neighbours (the assembly's strided slot loop; the per-slice SpMV
    and scaling loops instead of the unrolled uniform widths)."""
    p, vel = _hub_basin()
    say(f"high-valence mesh: {p.nnodes} nodes, {p.nelems2d} triangles")
    st = (p.head0, vel, p.head0, vel, p.head0, vel)
    o, g = pair(p, st)
    rp, _ = g.pattern()
    assert np.diff(rp).max() == 20
    o.shift(); g.shift()
    o.init(); g.init()
    assert nmismatch(o.resid(), g.resid()) == 0
    d_o, f_o = o.load()
    d_g, f_g, rr_g = g.load()
    assert nmismatch(d_o, d_g) == 0 and nmismatch(f_o, f_g) == 0
    assert bits_equal(o.residual(), rr_g)
    em, _ = g.elem_jacobians()
    for ie in range(p.nelems2d):
        assert bits_equal(o.elem_jacobian(ie), em[81 * ie: 81 * ie + 81])
    assert bits_equal(o.linear_setup(oopts(prec=2)), g.linear_sys_setup())
    o, g = pair(p, st)
    o.shift(); g.shift(); o.init(); g.init()
    r0 = np.abs(o.resid()).max()
    tol = 1e-8 * r0
    g.set_opts(default_opts(tol_nonlin=tol, max_nonlin_it=20))
    res_o = o.newton(oopts(tol_nonlin=tol, prec=2, max_nonlin_it=20))
    res_g = g.newton()
    say(f"  newton its oracle={res_o['iterations']} gpu={res_g['iterations']}")
    assert res_g["verdict"] == YES and res_o["ok"] == 1
    assert res_g["iterations"] == res_o["iterations"]
    assert np.allclose(res_g["hist_max"], res_o["hist_max"], rtol=0, atol=1e-10 * r0)
    for a, b in zip(o.get_state(), g.get_state()):
        assert np.abs(a - b).max() <= 1e-10 * max(1.0, np.abs(a).max())


def test_refined_element_blocks_are_fp32_roundings(stepped):
    """The refined forms store each FD element block as the FP32 rounding of
    the FP64 block that the FP64 form stores (the same FP64 arithmetic, then
    one rounding), and the assembly sums them in FP64: the assembled matrix
    stays within FP32 rounding of the FP64 form's."""
    name = "bumpy-60x30"
    p = CASES[name]
    say(f"refined element blocks: {name}")
    out = {}
    for prec in (0, 1):
        g = SW2Model(p, default_opts(precision=prec))
        g.set_state(*stepped[name])
        g.shift(); g.init(); g.resid()
        d, f, _ = g.load()
        em2, em1 = g.elem_jacobians()
        out[prec] = (d, f, em2, em1)
        g.close()
    (d0, f0, e0, b0), (d1, f1, e1, b1) = out[0], out[1]
    assert bits_equal(e1, e0.astype(np.float32).astype(np.float64))
    assert bits_equal(b1, b0.astype(np.float32).astype(np.float64))
    big = max(np.abs(d0).max(), np.abs(f0).max())
    err = max(np.abs(d1 - d0).max(), np.abs(f1 - f0).max()) / big
    say(f"  assembled refined vs FP64: {err:.2e} of the largest entry")
    assert err <= 1e-6


def test_newton_fp32_refined():
    """FP32 matrix/preconditioner/Krylov with FP64 residual and update reaches
    the same FP64 nonlinear tolerance; state within 1e-7 relative of FP64."""
    p = CASES["bumpy-60x30"]
    say("newton fp32 vs fp64: bumpy-60x30")
    g64 = SW2Model(p)
    g32 = SW2Model(p, default_opts(precision=1))
    g64.shift(); g64.init()
    r0 = np.abs(g64.resid()).max()
    for g in (g64, g32):
        g.set_opts(default_opts(tol_nonlin=1e-8 * r0, max_nonlin_it=30, precision=g.opts.precision))
    r64 = g64.newton()
    g32.shift()
    r32 = g32.newton()
    say(f"  fp64 its={r64['iterations']} fp32 its={r32['iterations']} final {r64['hist_max'][-1]:.3e} "
        f"{r32['hist_max'][-1]:.3e}")
    assert r64["verdict"] == YES and r32["verdict"] == YES
    assert r32["hist_max"][-1] <= 1e-8 * r0
    s64, s32 = g64.get_state(), g32.get_state()
    assert np.abs(s64[0] - s32[0]).max() <= 1e-7 * np.abs(s64[0]).max()


@pytest.mark.parametrize("name", ["floodplain-20x10", "bumpy-60x30"])
def test_autotuned_launch_shapes_change_nothing(name):
    """The autotuner picks block and grid sizes of the element list kernels;
    a Newton solve with the picked shapes, and with deliberately odd ones,
    equals the default-shape solve bit for bit."""
    from adhv8_kraken_amd.sw2 import Tuning
    p = CASES[name]
    tol = 1e-5 if name.startswith("flood") else 5e-8
    runs = []
    for mode in ("default", "auto", "odd", "lanes", "cols", "staged", "group"):
        g = SW2Model(p, default_opts(tol_nonlin=tol, max_nonlin_it=30))
        if mode == "auto":
            chosen = g.autotune(reps=2)
            say(f"  {name} autotuned: {chosen}")
            for v in chosen.values():
                assert v["block"] in (64, 128, 256) and v["grid"] >= 1
        elif mode == "odd":
            t = Tuning()
            for k in range(4):
                t.block[k] = (64, 256, 64, 256)[k]
                t.grid[k] = (3, 1, 7, 2)[k]
            t.layout[0] = (g.get_tuning().layout[0] + 1) % 4   # other Jacobian layouts than the default
            t.layout[1] = (g.get_tuning().layout[1] + 3) % 4
            g.set_tuning(t)
        elif mode in ("lanes", "cols", "staged", "group"):
            # every Jacobian layout on both lists: a lane per (triangle, column,
            # sign); column-uniform waves; column-uniform waves with the inputs in
            # LDS; a wave per triangle group doing all nine columns from LDS
            t = g.get_tuning()
            t.layout[0] = t.layout[1] = ("lanes", "cols", "staged", "group").index(mode)
            g.set_tuning(t)
        r = g.timestep(max_retries=2)
        h, v, *_ = g.get_state()
        runs.append((r, h, v))
        g.close()
    for r, h, v in runs[1:]:
        assert r["iterations"] == runs[0][0]["iterations"]
        assert np.array_equal(r["hist_max"], runs[0][0]["hist_max"])
        assert bits_equal(h, runs[0][1]) and bits_equal(v, runs[0][2])
