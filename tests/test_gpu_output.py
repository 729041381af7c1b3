"""Checkpoint / resume through a precise hot-start file on the device: the
library continues bit for bit, and ends where the oracle ends."""
import numpy as np
import pytest

from gpu_helpers import say
from adhv8_kraken_amd import deck, mesh, output
from adhv8_kraken_amd.sw2 import SW2Model, default_opts, YES

pytestmark = pytest.mark.gpu


def test_library_checkpoint_resume_bit_for_bit(tmp_path):
    p = mesh.bumpy_channel(40, 20, seed=6)
    g = SW2Model(p)
    g.shift(); g.init()
    tol = 1e-8 * float(np.abs(g.resid()).max())
    o = default_opts(tol_nonlin=tol, max_nonlin_it=30)
    g.close()
    say("checkpoint/resume: bumpy 40x20, 4 steps vs 2 + hot start + 2")

    def run(m, k):
        for _ in range(k):
            assert m.timestep()["verdict"] == YES

    a = SW2Model(p, o)
    run(a, 4)
    straight = a.get_state()
    a.close()
    b = SW2Model(p, o)
    run(b, 2)
    h, v, oh, ov, *_ = b.get_state()
    b.close()
    path = str(tmp_path / "ck.hot")
    output.write_hotstart(path, p.nnodes, p.nelems2d, 10.0, h, v, oh, ov, precise=True)
    hot = deck.read_hot(path, p.nnodes)
    c = SW2Model(p, o)
    vel = hot["IV"][:, :2].ravel()
    c.set_state(hot["IOH"], vel, hot["IPH"], hot["IPV"][:, :2].ravel(), hot["IOH"], vel)
    run(c, 2)
    resumed = c.get_state()
    c.close()
    assert np.array_equal(straight[0], resumed[0]) and np.array_equal(straight[1], resumed[1])
