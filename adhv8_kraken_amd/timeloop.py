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


def run_ensemble(ens, tf, dt_max, wall_s=None, max_retries=3, on_call=None):
    """The same time control for every model of an ensemble (`SW2Ensemble`),
    each with its own time and dt: the models not yet at tf take one step
    together per call (up to max_retries + 1 attempts, each failure cutting
    that model's dt by DT_REDUCE_FACTOR in the library, src/tools/tc_scale.c);
    a model whose attempts all failed stays where it was with the cut dt,
    which is how the reference retries until fe_main succeeds
    (src/main/adh_func.c:535-549); after a success its time advances and dt
    doubles, capped at dt_max and landed on tf (tc_end.c:5-50).  Stops when
    every model reached tf or after wall_s seconds.  Returns per-model times,
    attempt and failure counts, Newton iterations and the smallest dt used."""
    import time
    n = len(ens)
    t = np.zeros(n)
    dt = np.full(n, float(dt_max))
    for i in range(n):
        dt[i] = _land_on_final(0.0, dt[i], tf)
        ens.set_dt(i, dt[i])
    attempts = np.zeros(n, dtype=np.int64)
    failed = np.zeros(n, dtype=np.int64)
    newton = np.zeros(n, dtype=np.int64)
    min_dt = float(dt_max)
    calls = 0
    t0 = time.perf_counter()
    while (t < tf).any():
        if wall_s is not None and time.perf_counter() - t0 > wall_s:
            break
        which = (t < tf).astype(np.int32)
        rs = ens.timestep(max_retries=max_retries, which=which)
        calls += 1
        for i in np.nonzero(which)[0]:
            r = rs[i]
            newton[i] += r["iterations"]
            attempts[i] += len(r["hist_max"]) - r["iterations"]
            used = ens.get_dt(i)
            min_dt = min(min_dt, used)
            if r["verdict"] == 1:
                t[i] = tf if t[i] + used >= tf else t[i] + used
                if t[i] < tf:
                    nd = min(used * DT_ENLARGE_FACTOR, dt_max)
                    ens.set_dt(i, _land_on_final(t[i], nd, tf))
            else:
                failed[i] += 1
        if on_call is not None:
            on_call(calls, t, failed)
    return dict(t=t, attempts=attempts, failed_calls=failed, newton=newton, min_dt=min_dt, calls=calls,
                seconds=time.perf_counter() - t0)
