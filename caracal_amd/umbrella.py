"""Umbrella sampling along the reaction coordinate on the engine: calc_rate's
start-structure generation and window samplings (src/calc_rate.f90:650-1700).

Windows: n_samplings = nint((1 - lo)/dist) below xi = 1 and
n_over = nint((hi - 1)/dist) from xi = 1 up (src/calc_rate.f90:655-658), in
ascending order xi = 1 - n_samplings dist, ..., 1 - dist, 1, ..., 1 + (n_over - 1) dist.

Structure generation (src/calc_rate.f90:856-1098) is one trajectory walked
through the windows: from the transition state up through the windows above
xi = 1, then from the xi = 1 window's configuration down to the lowest
window, gen_step biased steps per window.  Bead 1 of each window's last
configuration, centred on its centre of mass, is that window's start
structure.  Before the walk down, the reference resets bead 1 only
(src/calc_rate.f90:966-972) and leaves the other beads where the highest
window left them; with widely spaced windows above xi = 1 that ring polymer's
centroid is further than xi_tol from the first window below, and every
attempt fails the check.  Here all beads return to the xi = 1 window's
configuration.

Window sampling (src/calc_rate.f90:1380-1700), for every window at once,
umbr_traj times: mdinit with the bias, equi_step steps, unbiased forces, then
umbr_step steps whose xi (mode 1) gives the trajectory's mean and variance.
The reference runs the umbr_traj trajectories of a window one after the other,
each from where the previous one ended; parallel_trajs=True runs them at the
same time from the window's start structure instead.  A trajectory that fails
the check starts again from its window's start structure; after more than
err_act_max (10) failures in a row the window takes the start structure of a
neighbour instead, one to the right, one to the left, two to the right, two
to the left (src/rpmd_check.f90:176-250, clamped to the windows that exist).
One whose variance exceeds 1e-2 is run again from where it ended, as the
reference does.

With histogram=(nbins, xi_min, bin_size) the sampling steps of every
trajectory are also counted into bins of xi on the device (bin j at
xi_min + j bin_size, the coordinate src/wham.f90 gives bin j), summed per
window over its umbr_traj accepted trajectories: the input of WHAM
(caracal_amd.rate.wham), which the reference reads from files that its own
sampling loop does not write (src/calc_rate.f90:1712-1724, :1808-1825).
"""
import math

import numpy as np

from . import _abi
from .engine import RPMDEngine

_DRAW_BASE = 1 << 41
ERR_ACT_MAX = 10                     # src/calc_rate.f90:219
_NEIGHBOUR = (0, 1, -1, 2, -2)       # start-structure offset after 0..4 resets


def _nint(x):
    """Fortran NINT: round half away from zero."""
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


class UmbrellaError(RuntimeError):
    """More failed trajectories than err_max (the reference stops then)."""


class UmbrellaSampling:
    """beta: scalar.  check: None or a dict with energy_tol (hartree per bead),
    xi_tol, err_max and ts_energy (hartree); rpmd_check's conditions
    (src/rpmd_check.f90:74-112)."""

    def __init__(self, surface, nbeads, dt, masses, reaction, beta, *, umbr_lo, umbr_hi, umbr_dist, k_force,
                 gen_step, equi_step, umbr_step, umbr_traj, andersen_step, check=None, precision=_abi.CAM_PREC_FP64,
                 seed=0, parallel_trajs=False, engine_factory=RPMDEngine, histogram=None):
        self.surface, self.nbeads, self.dt = surface, nbeads, dt
        self.masses = np.asarray(masses, dtype=np.float64)
        self.reaction, self.beta = reaction, float(beta)
        self.lo, self.hi, self.dist, self.k_force = umbr_lo, umbr_hi, umbr_dist, k_force
        self.gen_step, self.equi_step, self.umbr_step, self.umbr_traj = gen_step, equi_step, umbr_step, umbr_traj
        self.andersen_step, self.precision, self.seed = andersen_step, precision, seed
        self.parallel_trajs, self.check = parallel_trajs, check
        self._factory = engine_factory
        self.histogram = None if histogram is None else (int(histogram[0]), float(histogram[1]),
                                                          float(histogram[2]))
        self.histograms = None            # (windows, nbins) after sample() when histogram is set
        self._engines = {}
        self._created = 0
        self._draw = _DRAW_BASE
        self.errors = 0

    # -- windows, engines, draws ---------------------------------------------
    def windows(self):
        n_over = _nint((self.hi - 1.0) / self.dist)
        n_samp = _nint((1.0 - self.lo) / self.dist)
        below = [1.0 - i * self.dist for i in range(n_samp, 0, -1)]
        above = [1.0 + (i - 1) * self.dist for i in range(1, n_over + 1)]
        return np.array(below + above), n_samp

    def _engine(self, ntraj):
        if ntraj not in self._engines:
            self._engines[ntraj] = self._factory(
                self.surface, self.nbeads, ntraj, self.dt, self.masses, precision=self.precision,
                ensemble=_abi.CAM_ENS_UMBRELLA, reaction=self.reaction, andersen_step=self.andersen_step,
                seed=self.seed + self._created, remove_com=True)
            if self.histogram is not None:
                self._engines[ntraj].set_umbrella_histogram(*self.histogram)
            self._created += 1
        return self._engines[ntraj]

    def _next_draw(self):
        self._draw += 1
        return self._draw

    def _check_args(self, check_after):
        if self.check is None:
            return dict(check_after=-1)
        c = self.check
        return dict(epot_max=(c["ts_energy"] + c["energy_tol"]) * self.nbeads, xi_tol=c["xi_tol"],
                    check_after=check_after)

    def _count_errors(self, n):
        self.errors += n
        if self.check is not None and self.errors > self.check["err_max"]:
            raise UmbrellaError(f"{self.errors} failed trajectories (err_max {self.check['err_max']})")

    def close(self):
        for e in self._engines.values():
            e.close()
        self._engines.clear()

    # -- structure generation ---------------------------------------------------
    def _walk(self, q, xi, check_after):
        """gen_step biased steps of the walking trajectory in window xi from
        q, again from q after a failed check.  Returns the final positions."""
        e = self._engine(1)
        while True:
            e.set_windows([xi], [self.k_force])
            e.set_state(q[None], np.zeros_like(q)[None], [self.beta])
            e.draw_momenta(self._next_draw())
            e.begin_umbrella(bias=True, **self._check_args(check_after))
            e.run(self.gen_step)
            if not e.observables()["flags"][0] & 4:
                return e.get_state()[0][0]
            self._count_errors(1)

    def generate(self, ts_geometry):
        """Start structures (nwin, natoms, 3) of all windows, from the
        transition-state geometry (natoms, 3)."""
        xi_wins, n_samp = self.windows()
        struc = np.empty((len(xi_wins),) + np.shape(ts_geometry))
        q = np.broadcast_to(np.asarray(ts_geometry, dtype=np.float64), (self.nbeads,) + np.shape(ts_geometry)).copy()
        for idx in range(n_samp, len(xi_wins)):           # xi = 1 and up (checked after step 10)
            q = self._walk(q, xi_wins[idx], 10)
            struc[idx] = q[0]
            if idx == n_samp:
                q_top = q.copy()
        q = q_top                                          # back to the xi = 1 configuration
        for idx in range(n_samp - 1, -1, -1):              # below xi = 1, downwards
            q = self._walk(q, xi_wins[idx], 0)
            struc[idx] = q[0]
        m = self.masses[None, :, None]
        return struc - (struc * m).sum(axis=1, keepdims=True) / self.masses.sum()

    # -- window sampling -----------------------------------------------------------
    def _restart_structure(self, win, struc):
        """Count a failed check of window win and return the structure its
        trajectory starts again from."""
        self._err_act[win] += 1
        if self._err_act[win] > ERR_ACT_MAX:
            self._err_act[win] = 0
            self._resets[win] = min(self._resets[win] + 1, len(_NEIGHBOUR) - 1)
        k = min(max(win + _NEIGHBOUR[self._resets[win]], 0), len(struc) - 1)
        return struc[k]

    def _trajectories(self, q_start, xi, win, struc):
        """One trajectory per entry (window win[i] at xi[i]): mdinit,
        equilibration, unbiased start, sampling; failed ones again (see module
        doc).  Returns the final positions and the mean and variance of xi of
        each, and their histograms of xi (None without histogram)."""
        n = len(q_start)
        q_out, avg, var = np.empty_like(q_start), np.empty(n), np.empty(n)
        hist = None if self.histogram is None else np.zeros((n, self.histogram[0]))
        q_run = q_start.copy()
        pending = np.arange(n)
        while len(pending):
            e = self._engine(len(pending))
            e.set_windows(xi[pending], np.full(len(pending), self.k_force))
            e.set_state(q_run[pending], np.zeros_like(q_run[pending]), np.full(len(pending), self.beta))
            e.draw_momenta(self._next_draw())
            e.begin_umbrella(bias=True, **self._check_args(0))
            e.run(self.equi_step)
            bad = (e.observables()["flags"] & 4) != 0
            e.begin_umbrella(bias=False, **self._check_args(0))
            e.run(self.umbr_step)
            bad |= (e.observables()["flags"] & 4) != 0
            a, v = e.umbrella_stats()
            q_end, _ = e.get_state()
            noisy = ~bad & (v > 1e-2)
            ok = ~bad & ~noisy
            q_out[pending[ok]], avg[pending[ok]], var[pending[ok]] = q_end[ok], a[ok], v[ok]
            if hist is not None:
                hist[pending[ok]] = e.umbrella_histogram()[ok]
            for i in pending[ok]:                            # a finished trajectory clears its window's record
                self._err_act[win[i]] = self._resets[win[i]] = 0
            for i in pending[bad]:                           # check failed: a start structure again
                q_run[i] = self._restart_structure(win[i], struc)
            q_run[pending[noisy]] = q_end[noisy]             # variance too high: again from where it ended
            self._count_errors(int(bad.sum() + noisy.sum()))
            pending = pending[~ok]
        return q_out, avg, var, hist

    def sample(self, struc_equi):
        """Mean and variance of xi per window, averaged over its umbr_traj
        trajectories (src/calc_rate.f90:1692-1703); with histogram, the
        windows' histograms summed over those trajectories in
        self.histograms."""
        xi_wins, _ = self.windows()
        nw, nb = len(xi_wins), self.nbeads
        struc = np.repeat(np.asarray(struc_equi, dtype=np.float64)[:, None], nb, axis=1)
        self._err_act, self._resets = np.zeros(nw, dtype=int), np.zeros(nw, dtype=int)
        if self.parallel_trajs:
            rep = np.repeat(np.arange(nw), self.umbr_traj)
            _, a, v, h = self._trajectories(struc[rep], xi_wins[rep], rep, struc)
            if h is not None:
                self.histograms = h.reshape(nw, self.umbr_traj, -1).sum(axis=1)
            return xi_wins, a.reshape(nw, -1).mean(axis=1), v.reshape(nw, -1).mean(axis=1)
        average, variance = np.zeros(nw), np.zeros(nw)
        hsum = None if self.histogram is None else np.zeros((nw, self.histogram[0]))
        q = struc.copy()
        wins = np.arange(nw)
        for _ in range(self.umbr_traj):
            q, a, v, h = self._trajectories(q, xi_wins, wins, struc)
            average += a
            variance += v
            if h is not None:
                hsum += h
        self.histograms = hsum
        return xi_wins, average / self.umbr_traj, variance / self.umbr_traj
