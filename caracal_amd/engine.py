"""Host mirror of the reference's RPMD step driver (verlet, src/verlet.f90:65)
over the engine entry points of the C ABI."""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from ._abi import check, lib

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _ptr(a):
    return a.ctypes.data_as(_dp)


@dataclass
class Reaction:
    """Bimolecular reaction coordinate, keys of the reference's MECHA section
    (1-based atom indices, bohr)."""
    bond_form: list = field(default_factory=list)     # [(i, j), ...]
    bond_break: list = field(default_factory=list)
    form_ref: list = field(default_factory=list)
    break_ref: list = field(default_factory=list)
    reactants: list = field(default_factory=list)     # [[atoms of reactant 1], [atoms of reactant 2]]
    r_inf: float = 16.0
    xi_ideal: float = 1.0

    def to_c(self):
        r = _abi.Reaction()
        r.n_form, r.n_break = len(self.bond_form), len(self.bond_break)
        for i, (a, b) in enumerate(self.bond_form):
            r.bond_form[i][0], r.bond_form[i][1] = a, b
            r.form_ref[i] = self.form_ref[i]
        for i, (a, b) in enumerate(self.bond_break):
            r.bond_break[i][0], r.bond_break[i][1] = a, b
            r.break_ref[i] = self.break_ref[i]
        r.n_reac = len(self.reactants)
        for k, atoms in enumerate(self.reactants):
            r.reac_natoms[k] = len(atoms)
            for i, a in enumerate(atoms):
                r.reac_atoms[k][i] = a
        r.r_inf, r.xi_ideal = self.r_inf, self.xi_ideal
        return r


def make_config(surface, nbeads, ntraj, dt, masses, *, precision=_abi.CAM_PREC_FP64, nm_mode=_abi.CAM_NM_REFERENCE,
                ensemble=_abi.CAM_ENS_NVE, remove_com=False, andersen_step=0, steps_per_launch=0, seed=0,
                reaction=None, nm_kernel=_abi.CAM_NMK_AUTO, waves_per_eu=0, schedule=_abi.CAM_SCHED_AUTO, nose_q=0.0):
    """cam_config of one engine (or of one bucket of a batch)."""
    natoms = _abi.NATOMS[surface]
    masses = np.asarray(masses, dtype=np.float64)
    if masses.shape != (natoms,):
        raise ValueError(f"need {natoms} masses")
    c = _abi.Config()
    c.surface, c.natoms, c.nbeads, c.ntraj = surface, natoms, nbeads, ntraj
    c.precision, c.nm_mode, c.ensemble = precision, nm_mode, ensemble
    c.remove_com, c.andersen_step, c.steps_per_launch = int(remove_com), andersen_step, steps_per_launch
    c.nm_kernel = nm_kernel
    c.waves_per_eu = waves_per_eu
    c.schedule = schedule
    c.dt = dt
    for i, m in enumerate(masses):
        c.mass[i] = m
    c.seed = seed
    if reaction is not None:
        c.reaction = reaction.to_c()
    c.nose_q = nose_q
    return c


class RPMDEngine:
    """A batch of ntraj ring polymers of nbeads beads on one analytic surface,
    resident on the GPU.  State arrays have shape (ntraj, nbeads, natoms, 3)."""

    def __init__(self, surface, nbeads, ntraj, dt, masses, *, precision=_abi.CAM_PREC_FP64,
                 nm_mode=_abi.CAM_NM_REFERENCE, ensemble=_abi.CAM_ENS_NVE, remove_com=False,
                 andersen_step=0, steps_per_launch=0, seed=0, reaction=None, nm_kernel=_abi.CAM_NMK_AUTO,
                 waves_per_eu=0, schedule=_abi.CAM_SCHED_AUTO, nose_q=0.0):
        c = make_config(surface, nbeads, ntraj, dt, masses, precision=precision, nm_mode=nm_mode,
                        ensemble=ensemble, remove_com=remove_com, andersen_step=andersen_step,
                        steps_per_launch=steps_per_launch, seed=seed, reaction=reaction, nm_kernel=nm_kernel,
                        waves_per_eu=waves_per_eu, schedule=schedule, nose_q=nose_q)
        h = ctypes.c_void_p()
        check("cam_engine_create", lib().cam_engine_create(ctypes.byref(c), ctypes.byref(h)))
        self._attach(c, h, owned=True)

    def _attach(self, cfg, handle, owned):
        self.cfg = cfg
        self.natoms, self.nbeads, self.ntraj = cfg.natoms, cfg.nbeads, cfg.ntraj
        self._h = handle
        self._owned = owned
        self._hbins = 0

    @classmethod
    def _borrowed(cls, cfg, handle):
        """An engine owned by a batch (closing it is a no-op)."""
        e = cls.__new__(cls)
        e._attach(cfg, handle, owned=False)
        return e

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value and getattr(self, "_owned", False):
            lib().cam_engine_destroy(self._h)
        self._h = None

    def __del__(self):
        # at interpreter exit the library handle may already be gone
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def shape(self):
        return (self.ntraj, self.nbeads, self.natoms, 3)

    def set_state(self, q, p, beta=None):
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(self.shape)
        p = np.ascontiguousarray(p, dtype=np.float64).reshape(self.shape)
        bp = None
        if beta is not None:
            b = np.ascontiguousarray(np.broadcast_to(beta, (self.ntraj,)), dtype=np.float64)
            self._beta = b
            bp = _ptr(b)
        check("cam_engine_set_state", lib().cam_engine_set_state(self._h, _ptr(q), _ptr(p), bp))

    def get_state(self):
        q = np.empty(self.shape)
        p = np.empty(self.shape)
        check("cam_engine_get_state", lib().cam_engine_get_state(self._h, _ptr(q), _ptr(p)))
        return q, p

    def get_forces(self):
        g = np.empty(self.shape)
        v = np.empty((self.ntraj, self.nbeads))
        check("cam_engine_get_forces", lib().cam_engine_get_forces(self._h, _ptr(g), _ptr(v)))
        return v, g

    def draw_momenta(self, draw, antithetic_pairs=False):
        check("cam_engine_draw_momenta", lib().cam_engine_draw_momenta(self._h, draw, int(antithetic_pairs)))

    def constrain(self, project_p=True):
        st = np.zeros(self.ntraj, dtype=np.int32)
        check("cam_engine_constrain", lib().cam_engine_constrain(self._h, int(project_p), st.ctypes.data_as(_ip)))
        return st

    def begin_children(self):
        check("cam_engine_begin_children", lib().cam_engine_begin_children(self._h))

    def begin_parent(self, k_force=0.0, epot_max=float("inf"), check_after=-1):
        """Start recrossing parents (ensemble PARENT) as mdinit with bias_mode 2:
        fresh forces plus the umbrella bias of the centroid, and the first
        constraint direction.  Needed after every set_state.  Arms the
        per-step trajectory check from step check_after + 1 on (off if < 0);
        its result is bit 2 of observables()["flags"]."""
        check("cam_engine_begin_parent",
              lib().cam_engine_begin_parent(self._h, float(k_force), float(epot_max), int(check_after)))

    def set_windows(self, xi_ideal, k_force):
        """Ensemble UMBRELLA: window centre (xi, mode 1) and force constant of every trajectory."""
        x = np.ascontiguousarray(np.broadcast_to(xi_ideal, (self.ntraj,)), dtype=np.float64)
        k = np.ascontiguousarray(np.broadcast_to(k_force, (self.ntraj,)), dtype=np.float64)
        check("cam_engine_set_windows", lib().cam_engine_set_windows(self._h, _ptr(x), _ptr(k)))

    def begin_umbrella(self, bias=True, epot_max=float("inf"), xi_tol=float("inf"), check_after=-1):
        """Ensemble UMBRELLA: start a segment (bias=True: mdinit's biased
        forces; False: unbiased forces, as the sampling phase starts) and arm
        the trajectory check from step check_after + 1 on (off if < 0)."""
        check("cam_engine_begin_umbrella", lib().cam_engine_begin_umbrella(
            self._h, int(bool(bias)), float(epot_max), float(xi_tol), int(check_after)))

    def umbrella_stats(self):
        """Ensemble UMBRELLA: (average, variance) of xi per trajectory since begin_umbrella."""
        avg = np.empty(self.ntraj)
        var = np.empty(self.ntraj)
        check("cam_engine_umbrella_stats", lib().cam_engine_umbrella_stats(self._h, _ptr(avg), _ptr(var)))
        return avg, var

    def set_umbrella_histogram(self, nbins, xi_min, bin_size):
        """Ensemble UMBRELLA: count every step's xi into bin
        j = nint((xi - xi_min) / bin_size), 1 <= j <= nbins, per trajectory
        (nbins 0: off); cleared by begin_umbrella."""
        self._hbins = int(nbins)
        check("cam_engine_set_umbrella_histogram", lib().cam_engine_set_umbrella_histogram(
            self._h, int(nbins), float(xi_min), float(bin_size)))

    def umbrella_histogram(self):
        """(ntraj, nbins) float64 counts since begin_umbrella; column j - 1 is bin j."""
        h = np.empty((self.ntraj, max(self._hbins, 1)))
        check("cam_engine_umbrella_histogram", lib().cam_engine_umbrella_histogram(self._h, _ptr(h)))
        return h

    def run(self, nsteps, sync=True):
        check("cam_engine_run", lib().cam_engine_run(self._h, nsteps))
        if sync:
            self.sync()

    def prepare(self, nsteps):
        """Capture the HIP graph of an nsteps run without running it."""
        check("cam_engine_prepare", lib().cam_engine_prepare(self._h, nsteps))

    def sync(self):
        check("cam_engine_sync", lib().cam_engine_sync(self._h))

    def kappa(self, nsteps):
        num = np.empty(nsteps)
        den = np.empty(1)
        check("cam_engine_kappa", lib().cam_engine_kappa(self._h, _ptr(num), _ptr(den), nsteps))
        return num, den[0]

    def observables(self):
        n = self.ntraj
        ep, ek, es, xi = (np.empty(n) for _ in range(4))
        fl = np.empty(n, dtype=np.int32)
        check("cam_engine_observables",
              lib().cam_engine_observables(self._h, _ptr(ep), _ptr(ek), _ptr(es), _ptr(xi), fl.ctypes.data_as(_ip)))
        return dict(epot=ep, ekin=ek, espring=es, xi=xi, flags=fl)

    def autotune(self, trial_steps=100):
        """Time the candidate schedules on a copy of the state and keep the
        fastest; returns the trials as a list of dicts and the chosen one."""
        t = _abi.Tuning()
        check("cam_engine_autotune", lib().cam_engine_autotune(self._h, trial_steps, ctypes.byref(t)))
        trials = [dict(nm_kernel=t.nm_kernel[i], split=t.split[i], waves_per_eu=t.waves_per_eu[i],
                       steps_per_launch=t.steps_per_launch[i], us_per_step=t.us_per_step[i]) for i in range(t.n)]
        return trials, trials[t.chosen]

    def last_timing(self):
        ms = ctypes.c_double()
        spl = ctypes.c_int()
        nl = ctypes.c_int()
        check("cam_engine_last_timing",
              lib().cam_engine_last_timing(self._h, ctypes.byref(ms), ctypes.byref(spl), ctypes.byref(nl)))
        return ms.value, spl.value, nl.value


class RaggedBatch:
    """Trajectories of different surfaces and bead counts as one batch: one
    engine per bucket; a run is one captured graph with a branch per bucket.
    buckets: list of (surface, nbeads, dict(q, p, beta, masses, dt)) as
    workloads.ragged returns; every bucket gets the same ensemble options."""

    def __init__(self, buckets, **options):
        self.buckets = buckets
        cfgs = (_abi.Config * len(buckets))()
        for i, (surf, nb, d) in enumerate(buckets):
            cfgs[i] = make_config(surf, nb, len(d["beta"]), d["dt"], d["masses"], **options)
        h = ctypes.c_void_p()
        check("cam_batch_create", lib().cam_batch_create(cfgs, len(buckets), ctypes.byref(h)))
        self._h = h
        self.engines = []
        for i in range(len(buckets)):
            eh = ctypes.c_void_p()
            check("cam_batch_engine", lib().cam_batch_engine(h, i, ctypes.byref(eh)))
            self.engines.append(RPMDEngine._borrowed(cfgs[i], eh))

    def set_states(self):
        for e, (_, _, d) in zip(self.engines, self.buckets):
            e.set_state(d["q"], d["p"], d["beta"])

    def autotune(self, trial_steps=50):
        check("cam_batch_autotune", lib().cam_batch_autotune(self._h, trial_steps))

    def run(self, nsteps, sync=True):
        check("cam_batch_run", lib().cam_batch_run(self._h, nsteps))
        if sync:
            return self.sync()

    def sync(self):
        ms = ctypes.c_double()
        check("cam_batch_sync", lib().cam_batch_sync(self._h, ctypes.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            for e in self.engines:
                e._h = None
            lib().cam_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        # at interpreter exit the library handle may already be gone
        try:
            self.close()
        except Exception:
            pass
