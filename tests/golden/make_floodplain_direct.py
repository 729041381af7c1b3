"""Writes tests/golden/floodplain_direct_step0.npz: the oracle's first time
step of the smallest wet/dry floodplain of the 64-instance ensemble
(mesh.floodplain_ensemble: 224 x 112 cells, 50,176 triangles) with the
reference's default preconditioner, OP BLK 1 (one skyline LU of the whole
matrix), Newton to 1e-8 of the initial residual with at most 40 iterations.

The fixture holds the Newton iteration count, the residual history (max
and l2 norms) and the final depth and velocity at 512 evenly spaced nodes.
The oracle's OP BLK 1 factorisations take minutes on one core, which is why
the device test compares against this fixture instead of running the oracle.

Run from the repository root:
    python tests/golden/make_floodplain_direct.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from adhv8_kraken_amd import mesh  # noqa: E402
from oracle.oracle import Oracle, opts as oopts, set_log_mode  # noqa: E402


def smallest_instance():
    sizes = mesh.log_uniform_sizes(64, 50000, 500000, 5)
    k = int(np.argmin(sizes))
    ncx, ncy = mesh._grid_for(sizes[k])
    return k, ncx, ncy, mesh.floodplain(ncx, ncy, seed=5 * 1000 + k, dt=2.0)


def main():
    set_log_mode(1)
    k, ncx, ncy, p = smallest_instance()
    o = Oracle(p)
    o.shift(); o.init()
    r0 = float(np.abs(o.resid()).max())
    o.set_state(p.head0, p.vel0, p.head0, p.vel0, p.head0, p.vel0)
    o.shift(); o.init()
    res = o.newton(oopts(tol_nonlin=1e-8 * r0, prec=1, nblock=1, max_nonlin_it=40))
    h, v = o.get_state()[:2]
    idx = np.linspace(0, p.nnodes - 1, 512).astype(np.int64)
    v = np.asarray(v).reshape(-1, 2)
    out = os.path.join(HERE, "floodplain_direct_step0.npz")
    np.savez_compressed(out, instance=k, ncx=ncx, ncy=ncy, r0=r0, ok=res["ok"], iterations=res["iterations"],
                        hist_max=res["hist_max"], hist_l2=res["hist_l2"], nodes=idx, h=np.asarray(h)[idx],
                        u=v[idx, 0], v=v[idx, 1])
    print(f"instance {k} ({ncx}x{ncy}): ok {res['ok']} iterations {res['iterations']} "
          f"final {res['hist_max'][-1]:.3e} -> {out}")


if __name__ == "__main__":
    main()
