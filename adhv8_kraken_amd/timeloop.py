"""Host-side time control around the SW2 Newton step.

Mirrors the reference's single-model time loop (src/main/adh_func.c:505-549
with src/fe/fe_main.c:76-79, 353): every attempt shifts the state
(older <- old <- current, src/fe/sw2/fe_sw2.c:42-50) and runs Newton; a
failed Newton resets the iterate to the old state and cuts dt by 0.25
(src/tools/tc_scale.c, DT_REDUCE_FACTOR src/include/define.h:123) before the
attempt is repeated; after a success the time advances and the next dt is
taken from the DT series, doubled from the last one but capped by the series
and shortened to land on the final time (src/tools/tc_init.c:4-27,
src/tools/tc_end.c:5-50).

The model is any object with set_dt(dt), shift() and newton_ok() -> bool,
and optionally begin_step(t, dt), called before every attempt as the
reference updates its series averages and wind values there
(src/main/adh_func.c:538-546); `SW2Model` (HIP library) and the CPU oracle
wrapper used in tests both fit.
"""
from __future__ import annotations

import numpy as np

DT_REDUCE_FACTOR = 0.25
DT_ENLARGE_FACTOR = 2.0
SMALL_DT = 2.220446049250313e-16  # SMALL = DBL_EPSILON (src/include/define.h:27)


class TimeLoopError(RuntimeError):
    pass


def series_value(times, values, t):
    """Linear interpolation of a series at t, constant beyond its ends."""
    return float(np.interp(t, times, values))


def _land_on_final(t, dt, tf):
    if t + dt >= tf:
        return tf - t
    if t + 2.0 * dt >= tf:
        return 0.5 * (tf - t)
    return dt


def run(model, t0, tf, dt_times, dt_values, on_step=None, max_attempts=30):
    """Runs [t0, tf]; returns the list of (t, dt, attempts, newton_info)."""
    t = t0
    dt = _land_on_final(t, series_value(dt_times, dt_values, t), tf)
    if dt < SMALL_DT:
        raise TimeLoopError("time step is too small")
    log = []
    while True:
        attempts = 0
        while True:
            attempts += 1
            model.set_dt(dt)
            if hasattr(model, "begin_step"):
                model.begin_step(t, dt)
            model.shift()
            ok, info = model.newton_ok()
            if ok:
                break
            if attempts >= max_attempts:
                raise TimeLoopError(f"Newton failed {attempts} times at t={t}")
            dt *= DT_REDUCE_FACTOR
            if dt <= SMALL_DT:
                raise TimeLoopError("time step is ridiculous")
        log.append((t, dt, attempts, info))
        if on_step is not None:
            on_step(t, dt, info)
        t += dt
        if t >= tf:
            return log
        dt *= DT_ENLARGE_FACTOR
        dt_new = series_value(dt_times, dt_values, t)
        if dt_new < dt:
            dt = dt_new
        dt = _land_on_final(t, dt, tf)
