"""Host logic of the ensemble time control (adhv8_kraken_amd.timeloop.run_ensemble)
against a stand-in ensemble whose models converge only below a per-model dt
limit: retries cut dt by 0.25 inside a call (as the library does), a model
whose attempts all fail keeps its time and cut dt, successes double dt up to
the deck's dt and land on the end time (src/main/adh_func.c:535-549,
src/tools/tc_end.c:5-50).  CPU only."""
import numpy as np

from adhv8_kraken_amd import timeloop


class FakeEnsemble:
    def __init__(self, limits):
        self.limit = list(limits)
        self.dt = [0.0] * len(limits)
        self.log = []

    def __len__(self):
        return len(self.limit)

    def set_dt(self, i, dt):
        self.dt[i] = dt

    def get_dt(self, i):
        return self.dt[i]

    def timestep(self, max_retries=0, which=None):
        out = []
        for i in range(len(self)):
            if which is not None and not which[i]:
                out.append(dict(verdict=1, iterations=0, hist_max=np.zeros(0)))
                continue
            att = 0
            while True:
                att += 1
                if self.dt[i] <= self.limit[i]:
                    self.log.append((i, self.dt[i]))
                    out.append(dict(verdict=1, iterations=2, hist_max=np.zeros(att + 2)))
                    break
                self.dt[i] *= 0.25
                if att > max_retries:
                    out.append(dict(verdict=-3, iterations=0, hist_max=np.zeros(att)))
                    break
        return out


def test_every_model_reaches_the_end_time_with_cut_and_regrown_steps():
    ens = FakeEnsemble([10.0, 0.3, 0.001])
    res = timeloop.run_ensemble(ens, tf=7.0, dt_max=2.0, max_retries=1)
    assert np.all(res["t"] == 7.0)
    steps = {i: [dt for j, dt in ens.log if j == i] for i in range(3)}
    # model 0 never fails: 2, 2, 2, then the landing rule halves the rest (1 = 0.5 * (7 - 6))...
    assert abs(sum(steps[0]) - 7.0) < 1e-12 and max(steps[0]) == 2.0
    # model 1: every accepted dt is below its limit, and dt regrows after a cut
    assert all(d <= 0.3 for d in steps[1]) and abs(sum(steps[1]) - 7.0) < 1e-12
    # model 2 needs calls in which both attempts fail (2 -> 0.5 -> 0.125, ...), then advances
    assert res["failed_calls"][2] >= 1 and res["failed_calls"][0] == 0
    assert res["min_dt"] <= 0.001


def test_wall_cap_stops_early():
    ens = FakeEnsemble([-1.0])           # never converges
    res = timeloop.run_ensemble(ens, tf=1.0, dt_max=1.0, wall_s=0.05, max_retries=3)
    assert res["t"][0] == 0.0 and res["failed_calls"][0] >= 1
