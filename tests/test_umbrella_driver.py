"""The umbrella-sampling driver's orchestration (caracal_amd.umbrella) on the
CPU with a recording stand-in for the engine: calc_rate's windows
(src/calc_rate.f90:655-658), the structure walk up from the transition state
and then down from the xi = 1 window (:856-1098), and the window samplings
with their restarts (:1540-1700)."""
import numpy as np
import pytest

import caracal_amd as cam
from caracal_amd.umbrella import UmbrellaError, UmbrellaSampling


class FakeEngine:
    """run(n) moves the positions by n * 1e-3; begin_umbrella(bias=False)
    after a sampling run reports the window centre plus 0.001 as the mean."""
    log = []
    fail = []            # per equilibration run: trajectory indices whose check fails
    noisy = []           # per sampling run: trajectory indices whose variance is too high

    def __init__(self, surface, nbeads, ntraj, dt, masses, **kw):
        self.ntraj = ntraj
        # windows and the structure walk are verlet steps with constrain = 0,
        # which end with transrot (src/verlet.f90:1299-1305)
        assert kw.get("remove_com") is True
        self.log.append(("create", ntraj))

    def set_windows(self, xi, k):
        self.xi = np.array(xi, dtype=float)
        assert len(self.xi) == self.ntraj

    def set_state(self, q, p, beta):
        assert len(q) == self.ntraj
        self.q = np.array(q)
        self.log.append(("set_state", self.q[:, :, 0, 0].copy(), self.xi.copy()))

    def draw_momenta(self, draw):
        self.log.append(("draw", draw))

    def begin_umbrella(self, bias=True, **kw):
        self.phase = "equi" if bias else "samp"
        self.flags = np.zeros(self.ntraj, dtype=np.int32)
        self.log.append(("begin", bias, kw.get("check_after")))

    def run(self, n):
        self.q += n * 1e-3
        self.log.append(("run", self.phase, n))
        script = self.fail if self.phase == "equi" else self.noisy
        self.marked = script.pop(0) if script else []
        if self.phase == "equi":
            for i in self.marked:
                self.flags[i] = 4

    def observables(self):
        return dict(flags=self.flags)

    def umbrella_stats(self):
        var = np.full(self.ntraj, 1e-4)
        for i in self.marked:
            var[i] = 0.5
        return self.xi + 0.001, var

    def get_state(self):
        return self.q.copy(), np.zeros_like(self.q)

    def close(self):
        pass


def _us(**kw):
    FakeEngine.log, FakeEngine.fail, FakeEngine.noisy = [], [], []
    args = dict(umbr_lo=-0.05, umbr_hi=1.05, umbr_dist=0.01, k_force=0.05, gen_step=10, equi_step=4, umbr_step=6,
                umbr_traj=2, andersen_step=80, engine_factory=FakeEngine,
                check=dict(ts_energy=-0.17, energy_tol=0.1, xi_tol=0.1, err_max=5))
    args.update(kw)
    return UmbrellaSampling(cam.CAM_SURFACE_H3, 2, 1.0, [1.0, 1.0, 1.0], None, 1000.0, **args)


def test_windows_are_the_references():
    xi, n_samp = _us().windows()
    # nint(1.05/0.01) = 105 below xi = 1, nint(0.05/0.01) = 5 from xi = 1 up: 1.05 itself is not sampled
    assert n_samp == 105 and len(xi) == 110
    np.testing.assert_allclose(xi[[0, 104, 105, 109]], [-0.05, 0.99, 1.0, 1.04], atol=1e-12)
    assert (np.diff(xi) > 0).all()


def test_structure_walk_goes_up_then_down_from_the_xi_1_structure():
    us = _us(umbr_lo=0.97, umbr_hi=1.02)      # 3 windows below, 2 from 1 up
    xi, n_samp = us.windows()
    np.testing.assert_allclose(xi, [0.97, 0.98, 0.99, 1.0, 1.01], atol=1e-12)
    ts = np.zeros((3, 3))
    ts[:, 0] = [0.0, 1.0, 2.0]
    struc = us.generate(ts)
    sets = [e for e in FakeEngine.log if e[0] == "set_state"]
    assert [float(s[2][0]) for s in sets] == pytest.approx([1.0, 1.01, 0.99, 0.98, 0.97])
    # each window walks gen_step = 10 steps (+0.01) from where the last one ended; going down,
    # every bead starts from the xi = 1 window's configuration
    np.testing.assert_allclose(sets[0][1][0], [0.0, 0.0])
    np.testing.assert_allclose(sets[1][1][0], [0.01, 0.01])
    np.testing.assert_allclose(sets[2][1][0], [0.01, 0.01])
    np.testing.assert_allclose(sets[3][1][0], [0.02, 0.02])
    begins = [e for e in FakeEngine.log if e[0] == "begin"]
    assert [b[2] for b in begins] == [10, 10, 0, 0, 0]       # windows from xi = 1 up: checked after step 10
    # start structures: bead 1, centred (the fake moves every atom alike, so the centre of mass is removed)
    assert struc.shape == (5, 3, 3)
    np.testing.assert_allclose(struc.mean(axis=1), 0.0, atol=1e-12)


def test_window_sampling_chains_trajectories_and_restarts():
    us = _us(umbr_lo=0.98, umbr_hi=1.01)     # windows 0.98, 0.99, 1.0
    struc = np.zeros((3, 3, 3))
    struc[:, 0, 0] = [10.0, 20.0, 30.0]
    # trajectory 1: window 1 fails its check once, window 2 is noisy once
    FakeEngine.fail = [[1], [], []]
    FakeEngine.noisy = [[2], [], []]
    xi, avg, var = us.sample(struc)
    sets = [e for e in FakeEngine.log if e[0] == "set_state"]
    np.testing.assert_allclose(sets[0][1][:, 0], [10.0, 20.0, 30.0])
    # the check failure restarts from the window's start structure, the noisy one from where it ended
    np.testing.assert_allclose(sets[1][1][:, 0], [20.0, 30.01])
    # trajectory 2 of each window continues from where trajectory 1 ended
    np.testing.assert_allclose(sets[2][1][:, 0], [10.01, 20.01, 30.02])
    np.testing.assert_allclose(avg, xi + 0.001)
    np.testing.assert_allclose(var, 1e-4)
    assert us.errors == 2


def test_parallel_trajectories_and_error_limit():
    us = _us(umbr_lo=0.98, umbr_hi=1.01, parallel_trajs=True, umbr_traj=4)
    xi, avg, var = us.sample(np.zeros((3, 3, 3)))
    assert FakeEngine.log[0] == ("create", 12)
    np.testing.assert_allclose(avg, xi + 0.001)
    us = _us(umbr_lo=0.98, umbr_hi=1.01)
    FakeEngine.fail = [[0, 1, 2], [0, 1], [0, 1]]
    with pytest.raises(UmbrellaError):
        us.sample(np.zeros((3, 3, 3)))


def test_repeated_failures_take_a_neighbours_start_structure():
    us = _us(umbr_lo=0.97, umbr_hi=1.02, umbr_traj=1,
             check=dict(ts_energy=-0.17, energy_tol=0.1, xi_tol=0.1, err_max=100))
    struc = np.zeros((5, 3, 3))
    struc[:, 0, 0] = [10.0, 20.0, 30.0, 40.0, 50.0]
    # window 2 (xi = 0.99) fails 23 times: starts 2..11 from its own structure, 12..22 from window 3's,
    # then from window 1's
    FakeEngine.fail = [[2]] + [[0]] * 22
    us.sample(struc)
    starts = [e[1][:, 0] for e in FakeEngine.log if e[0] == "set_state"][1:]
    assert all(s[0] == 30.0 for s in starts[:10])
    assert all(s[0] == 40.0 for s in starts[10:21])
    assert starts[21][0] == 20.0 and starts[22][0] == 20.0
    assert us.errors == 23
