"""The recrossing driver's orchestration (caracal_amd.recross), on the CPU
with a recording stand-in for the engine: the order of mdinit, parent steps,
child batches and restarts of recross_serial (src/recross_serial.f90:100-305),
the Andersen schedule of the sampling cycles, and fresh random draws."""
import numpy as np
import pytest

import caracal_amd as cam
from caracal_amd.recross import Recrossing, RecrossingError


class FakeEngine:
    """Records the calls; run(n) moves every position by n * 1e-3 so that
    the positions tell how many steps a trajectory has made."""
    log = []
    fail_script = []          # per parent run: indices (within that run) whose check fails

    def __init__(self, surface, nbeads, ntraj, dt, masses, **kw):
        self.ntraj, self.kw = ntraj, kw
        self.log.append(("create", ntraj, kw["ensemble"], kw["andersen_step"], kw["seed"]))

    def set_state(self, q, p, beta):
        assert q.shape[0] == self.ntraj and len(beta) == self.ntraj
        self.q = np.array(q)
        self.log.append(("set_state", self.kw["ensemble"], self.q[:, 0, 0, 0].copy()))

    def draw_momenta(self, draw, antithetic_pairs=False):
        self.log.append(("draw", self.kw["ensemble"], draw, antithetic_pairs))

    def begin_parent(self, k_force=0.0, epot_max=float("inf"), check_after=-1):
        self.log.append(("begin_parent", k_force, epot_max, check_after))

    def begin_children(self):
        self.log.append(("begin_children",))

    def run(self, n):
        self.q += n * 1e-3
        self.log.append(("run", self.kw["ensemble"], n, self.kw["andersen_step"]))
        self.flags = np.zeros(self.ntraj, dtype=np.int32)
        if self.kw["ensemble"] == cam.CAM_ENS_PARENT and self.fail_script:
            for i in self.fail_script.pop(0):
                self.flags[i] = 4

    def get_state(self):
        return self.q.copy(), np.zeros_like(self.q)

    def observables(self):
        return dict(flags=self.flags)

    def kappa(self, nsteps):
        return np.arange(1.0, nsteps + 1), 8.0

    def close(self):
        pass


def _driver(**kw):
    FakeEngine.log = []
    args = dict(child_point=4, child_evol=6, child_interv=9, child_times=3, recr_equi=5, andersen_step=2,
                k_force=0.5, check=dict(ts_energy=-0.1, energy_tol=0.02, err_max=3), engine_factory=FakeEngine,
                seed=100)
    args.update(kw)
    return Recrossing(cam.CAM_SURFACE_H3, 4, 1.0, [1.0, 1.0, 1.0], None, 2.0, **args)


def test_sequence_follows_recross_serial():
    FakeEngine.fail_script = []
    r = _driver()
    q0 = np.zeros((3, 4, 3, 3)) + np.arange(3)[:, None, None, None]
    out = r.run(q0)
    log = FakeEngine.log
    kinds = [e[0] for e in log]
    # equilibration: one parent batch, mdinit (draw + begin) then recr_equi steps, checked after step 1000
    assert log[0] == ("create", 3, cam.CAM_ENS_PARENT, 2, 100)
    assert kinds[1:5] == ["set_state", "draw", "begin_parent", "run"]
    assert log[3] == ("begin_parent", 0.5, (-0.1 + 0.02) * 4, 1000)
    assert log[4] == ("run", cam.CAM_ENS_PARENT, 5, 2)
    # one child engine for all parents' children
    assert log[5] == ("create", 12, cam.CAM_ENS_CHILD, 0, 101)
    cyc = [i for i, e in enumerate(log) if e[0] == "begin_children"]
    assert len(cyc) == 3
    for c, i in enumerate(cyc, start=1):
        st, dr = log[i - 2], log[i - 1]
        # children start from the parents' configuration, four per parent, in +-p pairs
        assert st[0] == "set_state" and st[1] == cam.CAM_ENS_CHILD
        np.testing.assert_allclose(st[2], np.repeat(np.arange(3) + 0.005 + 0.009 * (c - 1), 4))
        assert dr[0] == "draw" and dr[3] is True
        assert log[i + 1] == ("run", cam.CAM_ENS_CHILD, 6, 0)
        # then the parent: mdinit and child_interv steps, checked from the first step;
        # Andersen at every step of cycle 3 (3 % int(sqrt(9)) == 0) and never in cycles 1, 2
        assert log[i + 2][0] == "create" or log[i + 2][0] == "set_state"
        run = next(e for e in log[i + 2:] if e[0] == "run")
        assert run == ("run", cam.CAM_ENS_PARENT, 9, 1 if c == 3 else 0)
        beg = next(e for e in log[i + 2:] if e[0] == "begin_parent")
        assert beg[3] == 0
    assert out["kappa"] == 6.0 / 8.0 and out["errors"] == 0
    np.testing.assert_allclose(out["q_parents"][:, 0, 0, 0], np.arange(3) + 0.005 + 3 * 0.009)
    # every draw is fresh, engines have distinct keys
    draws = [e[2] for e in log if e[0] == "draw"]
    assert len(set(draws)) == len(draws) == 1 + 3 + 3
    seeds = [e[4] for e in log if e[0] == "create"]
    assert len(set(seeds)) == len(seeds)


def test_interval_andersen_mode():
    FakeEngine.fail_script = []
    r = _driver(andersen="interval")
    r.run(np.zeros((2, 4, 3, 3)))
    runs = [e for e in FakeEngine.log if e[0] == "run" and e[1] == cam.CAM_ENS_PARENT]
    assert [e[3] for e in runs] == [2, 3, 3, 3]


def test_failed_parents_restart_from_their_start():
    # equilibration: parent 1 fails once; then in cycle 1 parent 0 fails twice (two restarts)
    FakeEngine.fail_script = [[1], [], [0], [0], []]
    r = _driver(child_times=1)
    q0 = np.zeros((3, 4, 3, 3)) + np.arange(3)[:, None, None, None]
    out = r.run(q0)
    log = FakeEngine.log
    sets = [e for e in log if e[0] == "set_state" and e[1] == cam.CAM_ENS_PARENT]
    np.testing.assert_allclose(sets[0][2], [0, 1, 2])
    np.testing.assert_allclose(sets[1][2], [1])                  # restart from the start structure
    np.testing.assert_allclose(sets[2][2], np.arange(3) + 0.005)  # cycle 1 from the configuration
    np.testing.assert_allclose(sets[3][2], [0.005])              # restart from the configuration
    np.testing.assert_allclose(sets[4][2], [0.005])
    assert out["errors"] == 3
    np.testing.assert_allclose(out["q_parents"][:, 0, 0, 0], np.arange(3) + 0.014)
    # a one-parent engine for the restarts, created once
    creates = [e for e in log if e[0] == "create"]
    assert [c[1] for c in creates] == [3, 1, 12, 3, 1]


def test_too_many_failures_stop_the_calculation():
    FakeEngine.fail_script = [[0, 1], [0], [0]]
    r = _driver(check=dict(ts_energy=0.0, energy_tol=0.0, err_max=3))
    with pytest.raises(RecrossingError):
        r.run(np.zeros((2, 4, 3, 3)))


def test_rejects_odd_child_point():
    with pytest.raises(ValueError):
        _driver(child_point=3)
