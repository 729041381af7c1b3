"""Shared set-up for the GPU parity tests: one problem, one oracle model and
one device model holding the same doubles."""
import sys

import numpy as np

from adhv8_kraken_amd import mesh
from adhv8_kraken_amd.sw2 import SW2Model, default_opts
from oracle.oracle import Oracle, opts as oopts, set_log_mode


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def stepped_problem(p, tol_rel=1e-6, max_it=8):
    """Advance the oracle one Newton step so velocities are non-zero, and
    return the problem with that state as its initial state."""
    set_log_mode(1)
    o = Oracle(p)
    o.shift()
    o.init()
    r0 = o.resid()
    o.newton(oopts(tol_nonlin=tol_rel * np.abs(r0).max(), max_nonlin_it=max_it, prec=2))
    h, v, oh, ov, ooh, oov = o.get_state()
    return h, v, oh, ov, ooh, oov


def pair(p, state=None, gopts=None):
    set_log_mode(1)
    o = Oracle(p)
    g = SW2Model(p, gopts or default_opts())
    if state is not None:
        o.set_state(*state)
        g.set_state(*state)
    return o, g


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def nmismatch(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return int(np.count_nonzero(a.view(np.int64) != b.view(np.int64)))
