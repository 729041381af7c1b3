"""adh_sw2_hip_init_noalloc (no device allocation) against adh_sw2_hip_init.

Two models in a row, the second smaller, each with Dirichlet node strings of
all three kinds (so the boundary mask is not trivially zero): the
allocation-free entry point returns the same mask, wet/dry flags and state as
the allocating one, bit for bit, although the second call reuses the staging
area the first one filled with more entries.
"""
import numpy as np
import pytest

from gpu_helpers import bits_equal, say
from adhv8_kraken_amd import mesh
from adhv8_kraken_amd.sw2 import SW2Model, default_opts, SW2Error

pytestmark = pytest.mark.gpu

BCT_PRS_DIR, BCT_VEL_DIR, BCT_VEL_PRS_DIR = 7, 8, 16


def _with_dirichlet(p, seed):
    rng = np.random.default_rng(seed)
    N = p.nnodes
    p.node_bc = np.zeros(N, dtype=np.int32)
    pick = rng.choice(N, size=N // 10, replace=False)
    p.node_bc[pick] = rng.choice([BCT_PRS_DIR, BCT_VEL_DIR, BCT_VEL_PRS_DIR], size=pick.size).astype(np.int32)
    p.node_bc_val = np.zeros(3 * N)
    p.node_bc_val[0::3] = rng.uniform(-0.1, 0.1, N)
    p.node_bc_val[1::3] = rng.uniform(-0.1, 0.1, N)
    p.node_bc_val[2::3] = rng.uniform(0.5, 1.5, N)
    return p


def test_init_noalloc_matches_init_on_two_calls_second_smaller():
    cases = [("floodplain-60x30", _with_dirichlet(mesh.floodplain(60, 30), 1)),
             ("channel-40x20", _with_dirichlet(mesh.channel_config1(), 2))]
    assert cases[1][1].nnodes < cases[0][1].nnodes
    for name, p in cases:
        say(f"init_noalloc: {name} ({p.nnodes} nodes)")
        outs = []
        for entry in ("init", "init_noalloc"):
            g = SW2Model(p, default_opts())
            g.shift()
            bc, wd = getattr(g, entry)()
            outs.append((bc, wd, g.get_state()[:2]))
            g.close()
        (bc0, wd0, s0), (bc1, wd1, s1) = outs
        assert np.array_equal(bc0, bc1) and np.array_equal(wd0, wd1)
        assert bits_equal(s0[0], s1[0]) and bits_equal(s0[1], s1[1])
        assert (bc0 == 1).sum() > 0 and set(np.unique(bc0)) <= {1, -3}   # YES = 1 / NO = -3


def test_init_noalloc_refuses_a_model_beyond_its_staging_area():
    p = mesh.bumpy_channel(2100, 1000)   # 2,103,101 nodes > 2,097,152
    g = SW2Model(p, default_opts())
    with pytest.raises(SW2Error, match="staging area"):
        g.init_noalloc()
    g.close()
