"""Recrossing factor kappa(t) on the engine: recross_serial
(src/recross_serial.f90:100-330) for a batch of independent parent
trajectories.

Each parent follows the reference's sequence:

1. mdinit (fresh forces, the umbrella bias of the centroid, momenta drawn;
   src/mdinit.f90:79-130) and recr_equi constrained parent steps with the
   Andersen interval int(sqrt(recr_equi)) (src/recross_serial.f90:91-136),
   checked from step 1001 on; a parent that fails the check starts again
   from its start structure (goto 112);
2. child_times cycles (:156-305): from the parent's configuration,
   child_point/2 pairs of children with fresh momenta +p and -p, each run
   child_evol steps while kappa's numerator and denominator accumulate; then
   the parent returns to that configuration, mdinit again, and child_interv
   constrained steps, checked every step, starting again from the
   configuration after a failed check (goto 113).

The children of all parents of one cycle run as one batch, and num(t) and den
accumulate over parents as the reference accumulates them over cycles.

The reference passes the cycle index, not the step, as verlet's istep in the
sampling loop (src/recross_serial.f90:286), so there a cycle whose index is a
multiple of int(sqrt(child_interv)) redraws the momenta at every step and the
other cycles never do.  andersen="reference" reproduces that;
andersen="interval" redraws every int(sqrt(child_interv)) steps instead.

Random numbers come from the engine's counter-based stream: each engine this
driver creates gets the key seed + (its creation index), and every mdinit or
child draw takes the next draw id from 2**40 up, so no two draws repeat.
"""
import math

import numpy as np

from . import _abi
from .engine import RPMDEngine

_DRAW_BASE = 1 << 40


class RecrossingError(RuntimeError):
    """More failed parent segments than err_max (rpmd_check stops the program
    then, src/rpmd_check.f90:149-156)."""


class Recrossing:
    """One recrossing calculation for nparents independent parents.

    beta: scalar or (nparents,).  andersen_step: the equilibration's Andersen
    interval, int(sqrt(recr_equi)) as the reference sets it unless given.
    check: None (recross_check off) or a dict with energy_tol (hartree per
    bead), err_max and ts_energy (hartree, the surface's energy at the
    transition state); a parent fails its check when its summed bead energy
    exceeds (ts_energy + energy_tol) * nbeads or its positions are not finite
    (src/rpmd_check.f90:74-98)."""

    def __init__(self, surface, nbeads, dt, masses, reaction, beta, *, child_point, child_evol, child_interv,
                 child_times, recr_equi, andersen_step=None, k_force=0.0, check=None, precision=_abi.CAM_PREC_FP64,
                 seed=0, andersen="reference", engine_factory=RPMDEngine):
        if child_point < 2 or child_point % 2:
            raise ValueError("child_point must be a positive even number (children come in +-p pairs)")
        if andersen not in ("reference", "interval"):
            raise ValueError("andersen must be 'reference' or 'interval'")
        for name, v in (("child_evol", child_evol), ("child_interv", child_interv), ("child_times", child_times),
                        ("recr_equi", recr_equi)):
            if v < 1:
                raise ValueError(f"{name} must be >= 1")
        self.surface, self.nbeads, self.dt = surface, nbeads, dt
        self.masses = np.asarray(masses, dtype=np.float64)
        self.reaction, self.beta0 = reaction, beta
        self.child_point, self.child_evol, self.child_interv = child_point, child_evol, child_interv
        # the reference resets andersen_step to int(sqrt(recr_equi)) for the equilibration
        self.child_times, self.recr_equi = child_times, recr_equi
        self.andersen_step = int(math.sqrt(recr_equi)) if andersen_step is None else andersen_step
        self.k_force, self.precision, self.seed, self.andersen = k_force, precision, seed, andersen
        self.check = check
        if check is not None:
            self.epot_max = (check["ts_energy"] + check["energy_tol"]) * nbeads
            self.err_max = check["err_max"]
        self._factory = engine_factory
        self._parents = {}
        self._created = 0
        self._draw = _DRAW_BASE
        self.errors = 0

    # -- engines and draws ------------------------------------------------
    def _engine(self, ntraj, ensemble, andersen_step=0):
        e = self._factory(self.surface, self.nbeads, ntraj, self.dt, self.masses, precision=self.precision,
                          ensemble=ensemble, reaction=self.reaction, andersen_step=andersen_step,
                          seed=self.seed + self._created)
        self._created += 1
        return e

    def _parent_engine(self, ntraj, andersen_step):
        key = (ntraj, andersen_step)
        if key not in self._parents:
            self._parents[key] = self._engine(ntraj, _abi.CAM_ENS_PARENT, andersen_step)
        return self._parents[key]

    def _next_draw(self):
        self._draw += 1
        return self._draw

    # -- parent segments ----------------------------------------------------
    def _segment(self, q_start, beta, nsteps, andersen_step, check_after):
        """mdinit and nsteps constrained steps for every parent from q_start;
        a parent whose check fails starts again from its own q_start until
        all pass.  Returns the parents' final positions."""
        out = np.empty_like(q_start)
        pending = np.arange(len(q_start))
        while len(pending):
            e = self._parent_engine(len(pending), andersen_step)
            e.set_state(q_start[pending], np.zeros_like(q_start[pending]), beta[pending])
            e.draw_momenta(self._next_draw())
            if self.check is None:
                e.begin_parent(self.k_force)
            else:
                e.begin_parent(self.k_force, epot_max=self.epot_max, check_after=check_after)
            e.run(nsteps)
            q, _ = e.get_state()
            failed = (e.observables()["flags"] & 4) != 0
            out[pending[~failed]] = q[~failed]
            if failed.any():
                self.errors += int(failed.sum())
                if self.errors > self.err_max:
                    raise RecrossingError(f"{self.errors} failed parent segments (err_max {self.err_max})")
            pending = pending[failed]
        return out

    def _sampling_andersen(self, cycle):
        interval = int(math.sqrt(self.child_interv))
        if self.andersen == "interval":
            return interval
        return 1 if cycle % interval == 0 else 0

    # -- the calculation ------------------------------------------------------
    def run(self, q0):
        """q0: (nparents, nbeads, natoms, 3) start structures.  Returns a dict
        with kappa = num(child_evol)/den, kappa_t (the recrossing factor at
        child steps 1..child_evol), num, den, the parents' final positions
        and the number of failed segments."""
        q0 = np.ascontiguousarray(q0, dtype=np.float64)
        n = len(q0)
        beta = np.ascontiguousarray(np.broadcast_to(self.beta0, (n,)), dtype=np.float64)
        q = self._segment(q0, beta, self.recr_equi, self.andersen_step, 1000)
        child = self._engine(n * self.child_point, _abi.CAM_ENS_CHILD)
        beta_c = np.repeat(beta, self.child_point)
        try:
            for cycle in range(1, self.child_times + 1):
                qc = np.repeat(q, self.child_point, axis=0)
                child.set_state(qc, np.zeros_like(qc), beta_c)
                child.draw_momenta(self._next_draw(), antithetic_pairs=True)
                child.begin_children()
                child.run(self.child_evol)
                q = self._segment(q, beta, self.child_interv, self._sampling_andersen(cycle), 0)
            num, den = child.kappa(self.child_evol)
        finally:
            child.close()
            for e in self._parents.values():
                e.close()
            self._parents.clear()
        return dict(kappa=num[-1] / den, kappa_t=num / den, num=num, den=den, q_parents=q, errors=self.errors)
