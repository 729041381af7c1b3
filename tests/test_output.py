"""Solution output and hot-start files (SMS DATASET, the reference's
ssw_2d_print_ts / hot-start reader formats): layout, read-back, and a
checkpoint/resume round trip that continues bit for bit.  CPU only (oracle)."""
import os

import numpy as np

from adhv8_kraken_amd import deck, mesh, output
from oracle.oracle import Oracle, opts


def _run(o, steps, tol):
    for _ in range(steps):
        o.shift()
        r = o.newton(opts(tol_nonlin=tol, prec=1, nblock=1, max_nonlin_it=20))
        assert r["ok"] == 1


def test_solution_writer_layout_and_readback(tmp_path):
    p = mesh.channel_config1()
    rng = np.random.default_rng(1)
    h = rng.uniform(0.5, 1.5, p.nnodes)
    v = rng.uniform(-1, 1, 2 * p.nnodes)
    stem = str(tmp_path / "run")
    with output.SolutionWriter(stem, p.nnodes, p.nelems2d) as w:
        w.write(0.0, h, v, h, v)
        w.write(10.0, 2 * h, -v, h, v)
    lines = open(stem + "_dep.dat").read().splitlines()
    assert lines[:7] == ["DATASET", 'OBJTYPE "mesh2d"', "BEGSCL", f"ND {p.nnodes}", f"NC {p.nelems2d}",
                         'NAME "Depth"', "TIMEUNITS SECONDS"]
    assert lines[7] == "TS 0 %15.8e" % 0.0 and lines[-1] == "ENDDS"
    assert lines[8] == "%16.8e" % h[0]
    vl = open(stem + "_vel.dat").read().splitlines()
    assert vl[2] == "BEGVEC" and vl[8] == "%16.8e %16.8e %16.8e" % (v[0], v[1], 0.0)
    t, sets = output.read_dataset(stem + "_dep.dat")
    assert list(t) == [0.0, 10.0]
    np.testing.assert_allclose(sets[1], 2 * h, rtol=1e-8)      # 9 significant digits
    t, vs = output.read_dataset(stem + "_vel.dat")
    np.testing.assert_allclose(vs[1].ravel(), -v, rtol=1e-8, atol=1e-12)


def test_hotstart_is_read_by_the_deck_reader(tmp_path):
    p = mesh.channel_config1()
    rng = np.random.default_rng(2)
    h, oh = rng.uniform(0.5, 1.5, p.nnodes), rng.uniform(0.5, 1.5, p.nnodes)
    v, ov = rng.uniform(-1, 1, 2 * p.nnodes), rng.uniform(-1, 1, 2 * p.nnodes)
    path = str(tmp_path / "x.hot")
    output.write_hotstart(path, p.nnodes, p.nelems2d, 5.0, h, v, oh, ov, precise=True)
    hot = deck.read_hot(path, p.nnodes)
    assert np.array_equal(hot["IOH"], h) and np.array_equal(hot["IPH"], oh)
    assert np.array_equal(hot["IV"][:, :2].ravel(), v) and np.array_equal(hot["IPV"][:, :2].ravel(), ov)


def test_checkpoint_resume_continues_bit_for_bit(tmp_path):
    """4 steps straight through, against 2 steps, a precise hot start
    written and read back (older level = current, as the reference's
    hot-start reader sets it, ssw.c:510-542), and 2 more steps."""
    p = mesh.channel_config1()
    o = Oracle(p)
    o.shift(); o.init()
    tol = 1e-8 * float(np.abs(o.resid()).max())
    o.set_state(p.head0, p.vel0, p.head0, p.vel0, p.head0, p.vel0)
    _run(o, 4, tol)
    straight = o.get_state()

    a = Oracle(p)
    _run(a, 2, tol)
    h, v, oh, ov, *_ = a.get_state()
    path = str(tmp_path / "ck.hot")
    output.write_hotstart(path, p.nnodes, p.nelems2d, 20.0, h, v, oh, ov, precise=True)
    hot = deck.read_hot(path, p.nnodes)
    b = Oracle(p)
    b.set_state(hot["IOH"], hot["IV"][:, :2].ravel(), hot["IPH"], hot["IPV"][:, :2].ravel(), hot["IOH"], hot["IV"][:, :2].ravel())
    _run(b, 2, tol)
    resumed = b.get_state()
    for x, y in zip(straight[:2], resumed[:2]):
        assert np.array_equal(x, y)
