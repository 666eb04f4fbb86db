"""The RPMD rate constant of a bimolecular reaction, calc_rate's full
sequence (src/calc_rate.f90) on one GPU: start structures along the
reaction coordinate, umbrella windows, umbrella integration into the PMF,
the recrossing factor at the PMF maximum, and the QTST rate (calc_k_t).

The parameters keep the names and meaning of calc_rate's key-file sections
(umbrella, pmf, recross; examples/calc_rate/h+h2/rate.key is the template):

  umbrella: bias (hartree per kelvin: the force constant is bias * kelvin,
            src/calc_rate.f90:698-699), bonds (lo, hi), dist, gen_steps,
            equi_steps, sample_steps, sample_trajs
  pmf:      xi_range (xi_min, xi_max), bins, method (pmf_method: INTEGRATION,
            WHAM or ALL; WHAM takes the windows' histograms of xi, collected
            on the device; ALL computes both and keeps the integration's, as
            src/calc_rate.f90:2018-2021 does)
  recross:  equi_steps, child_total, child_interval, child_perpoint,
            child_steps, nparents (independent parents; 1 in the reference)
"""
import time
from dataclasses import dataclass

import numpy as np

from . import _abi, rate
from .engine import Reaction
from .recross import Recrossing
from .umbrella import UmbrellaSampling
from .units import beta_of

KJMOL_PER_HARTREE = 627.5094743 * 4.1840       # hartree * joule of the reference


@dataclass
class UmbrellaKeys:
    bias: float
    bonds: tuple
    dist: float
    gen_steps: int
    equi_steps: int
    sample_steps: int
    sample_trajs: int


@dataclass
class PmfKeys:
    xi_range: tuple
    bins: int
    method: str = "INTEGRATION"


@dataclass
class RecrossKeys:
    equi_steps: int
    child_total: int
    child_interval: int
    child_perpoint: int
    child_steps: int
    nparents: int = 1


def _with_xi_ideal(reaction, xi):
    return Reaction(bond_form=reaction.bond_form, bond_break=reaction.bond_break, form_ref=reaction.form_ref,
                    break_ref=reaction.break_ref, reactants=reaction.reactants, r_inf=reaction.r_inf, xi_ideal=xi)


class RPMDRate:
    """energy_tol in kJ/mol (the reference's unit; default 250), xi_tol
    (default 0.1) and err_max (default 1000) configure rpmd_check; check=False
    turns it off.  andersen_step is the nvt section's value."""

    def __init__(self, surface, nbeads, dt, masses, reaction, kelvin, *, umbrella, pmf, recross, andersen_step,
                 n_paths=1, check=True, energy_tol=250.0, xi_tol=0.1, err_max=1000, seed=0,
                 precision=_abi.CAM_PREC_FP64, parallel_trajs=False, pmf_minloc="ZERO"):
        if recross.child_total % recross.child_perpoint:
            raise ValueError("child_total must be a multiple of child_perpoint")
        self.surface, self.nbeads, self.dt, self.masses = surface, nbeads, dt, np.asarray(masses, dtype=np.float64)
        self.reaction, self.kelvin, self.beta = reaction, float(kelvin), float(beta_of(kelvin))
        self.k_force = umbrella.bias * self.kelvin
        self.um, self.pmf, self.rc = umbrella, pmf, recross
        self.andersen_step, self.n_paths, self.seed, self.precision = andersen_step, n_paths, seed, precision
        self.parallel_trajs, self.pmf_minloc = parallel_trajs, pmf_minloc
        self.check = dict(energy_tol=energy_tol / KJMOL_PER_HARTREE, xi_tol=xi_tol, err_max=err_max) if check else None

    def _mass_reac(self):
        return [float(sum(self.masses[a - 1] for a in atoms)) for atoms in self.reaction.reactants[:2]]

    def run(self, ts_geometry, ts_energy):
        """ts_geometry: (natoms, 3) bohr; ts_energy: the surface's energy
        there (hartree), the reference of rpmd_check's energy limit."""
        check = None if self.check is None else dict(self.check, ts_energy=ts_energy)
        um = self.um
        method = self.pmf.method.upper()
        if method not in ("INTEGRATION", "WHAM", "ALL"):
            raise ValueError("pmf method must be INTEGRATION, WHAM or ALL")
        xi_min, xi_max = self.pmf.xi_range
        bin_size = (xi_max - xi_min) / (self.pmf.bins - 1)         # src/calc_rate.f90:1803, :1949
        hist = None if method == "INTEGRATION" else (self.pmf.bins, xi_min, bin_size)
        us = UmbrellaSampling(self.surface, self.nbeads, self.dt, self.masses, self.reaction, self.beta,
                              umbr_lo=um.bonds[0], umbr_hi=um.bonds[1], umbr_dist=um.dist, k_force=self.k_force,
                              gen_step=um.gen_steps, equi_step=um.equi_steps, umbr_step=um.sample_steps,
                              umbr_traj=um.sample_trajs, andersen_step=self.andersen_step, check=check,
                              precision=self.precision, seed=self.seed, parallel_trajs=self.parallel_trajs,
                              histogram=hist)
        wall = {}
        t0 = time.perf_counter()
        try:
            struc = us.generate(ts_geometry)
            wall["structures"] = time.perf_counter() - t0
            xi_wins, average, variance = us.sample(struc)
            wall["windows"] = time.perf_counter() - t0 - wall["structures"]
        finally:
            us.close()
        dA, bin_coord, pmf = rate.umbrella_integration(xi_wins, average, variance, np.full(len(xi_wins), self.k_force),
                                                       self.beta, xi_min, xi_max, self.pmf.bins,
                                                       um.sample_trajs, um.sample_steps)
        extra = {}
        if hist is not None:
            # WHAM over the windows that hold samples (all of them here; the
            # reference's last window never has a histogram)
            hw = us.histograms
            keep = hw.sum(axis=1) > 0
            coord_w, pmf_w, _, it = rate.wham(hw[keep], xi_wins[keep], np.full(int(keep.sum()), self.k_force),
                                              xi_min, xi_max, self.pmf.bins, self.kelvin)
            extra = dict(wham_coord=coord_w, pmf_wham=pmf_w, wham_iterations=it, histograms=hw)
            if method == "WHAM":
                # the WHAM PMF in hartree (the reference stores it times
                # 627.509 x 0.52917721092, src/wham.f90:189, and then uses it
                # as hartree in the rate); its bins are centred at coord_w
                dA, bin_coord, pmf = None, coord_w, pmf_w
        imax, imin = rate.pmf_extrema(pmf, xi_min, xi_max, self.pmf_minloc)
        xi_barrier = float(bin_coord[imax])
        # the parent starts from the start structure of the window nearest the barrier, every bead alike
        ts_locate = int(np.argmin(np.abs(xi_wins - xi_barrier)))
        rc = self.rc
        q0 = np.broadcast_to(struc[ts_locate], (rc.nparents, self.nbeads) + struc.shape[1:]).copy()
        rec = Recrossing(self.surface, self.nbeads, self.dt, self.masses, _with_xi_ideal(self.reaction, xi_barrier),
                         self.beta, child_point=rc.child_perpoint, child_evol=rc.child_steps,
                         child_interv=rc.child_interval, child_times=rc.child_total // rc.child_perpoint,
                         recr_equi=rc.equi_steps, k_force=self.k_force, check=check, seed=self.seed + 1000,
                         precision=self.precision)
        t1 = time.perf_counter()
        kres = rec.run(q0)
        wall["recrossing"] = time.perf_counter() - t1
        kappa = kres["kappa"]
        if kappa < 0.002:                 # src/calc_rate.f90:2236-2239
            kappa = 1.0
        k_mol = rate.rate_bimolecular(kappa, self.beta, pmf[imax], pmf[imin], self._mass_reac(), self.reaction.r_inf,
                                      self.n_paths)
        return dict(xi_wins=xi_wins, average=average, variance=variance, start_structures=struc, dA=dA,
                    bin_coord=bin_coord, pmf=pmf if method == "WHAM" else pmf[:-1], xi_barrier=xi_barrier,
                    pmf_max=float(pmf[imax]), **extra,
                    pmf_min=float(pmf[imin]), kappa=kappa, kappa_t=kres["kappa_t"], k_cm3_mol_s=k_mol,
                    k_cm3_molecule_s=k_mol / rate.AVOGADRO, errors=us.errors + kres["errors"], wall_s=wall)


def h_h2_example(scale=1.0, kelvin=300.0, seed=0, pmf_method="INTEGRATION"):
    """examples/calc_rate/h+h2/rate.key: H + H2 on BKMP2 at 300 K with 8
    beads, dt 0.1 fs; every step count multiplied by scale; pmf_method as
    the key file's pmf section names it (the example's is INTEGRATION)."""
    from . import CAM_SURFACE_H3, egrad_h3, workloads
    from .units import FS_TO_AU, mass_au

    def n(x):
        return max(1, int(round(x * scale)))

    perpoint = 100
    r = RPMDRate(CAM_SURFACE_H3, 8, 0.1 * FS_TO_AU, mass_au([1.008] * 3), workloads.h3_reaction(), kelvin,
                 umbrella=UmbrellaKeys(bias=0.05, bonds=(-0.05, 1.05), dist=0.01, gen_steps=n(10000),
                                       equi_steps=n(10000), sample_steps=n(20000), sample_trajs=10),
                 pmf=PmfKeys(xi_range=(-0.05, 1.05), bins=5000, method=pmf_method),
                 recross=RecrossKeys(equi_steps=n(50000), child_total=perpoint * max(1, n(100)),
                                     child_interval=n(1000), child_perpoint=perpoint, child_steps=n(500)),
                 andersen_step=80, n_paths=2, seed=seed)
    ts = workloads.H3_TS
    return r.run(ts, float(egrad_h3(ts[None])[0][0]))


def h_h2_kappa(seed, nm_mode, scale=1.0, kelvin=300.0, xi_barrier=1.0):
    """The recrossing part of examples/calc_rate/h+h2/rate.key alone: parents
    started at the BKMP2 saddle, constrained to xi_barrier, equilibrated for
    50,000 steps, then 100 cycles of 100 children x 500 steps every 1000
    parent steps (step counts times scale).  nm_mode: the normal-mode basis
    of every engine (CAM_NM_REFERENCE, the reference's cosine transform, or
    CAM_NM_ORTHOGONAL).  Returns kappa and kappa(t)."""
    import functools

    from . import CAM_SURFACE_H3, RPMDEngine, workloads
    from .units import FS_TO_AU, beta_of, mass_au

    def n(x):
        return max(1, int(round(x * scale)))

    nb = 8
    factory = functools.partial(RPMDEngine, nm_mode=nm_mode)
    rec = Recrossing(CAM_SURFACE_H3, nb, 0.1 * FS_TO_AU, mass_au([1.008] * 3),
                     _with_xi_ideal(workloads.h3_reaction(), xi_barrier), float(beta_of(kelvin)), child_point=100,
                     child_evol=n(500), child_interv=n(1000), child_times=max(1, n(100)), recr_equi=n(50000),
                     seed=seed, engine_factory=factory)
    q0 = np.broadcast_to(workloads.H3_TS, (1, nb) + workloads.H3_TS.shape).copy()
    res = rec.run(q0)
    return dict(kappa=float(res["kappa"]), kappa_t=res["kappa_t"], errors=res["errors"])


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser(description="RPMD rate of H + H2 (the reference's calc_rate example) on one GPU")
    ap.add_argument("--scale", type=float, default=1.0, help="multiply every step count (1 = the key file's)")
    ap.add_argument("--kelvin", type=float, default=300.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--pmf-method", default="INTEGRATION", choices=["INTEGRATION", "WHAM", "ALL"])
    ap.add_argument("--kappa-seeds", type=int, default=0,
                    help="run only the recrossing at xi = 1 for this many seeds in both normal-mode bases")
    args = ap.parse_args()
    if args.kappa_seeds:
        from ._abi import CAM_NM_ORTHOGONAL, CAM_NM_REFERENCE
        out = {}
        for name, mode in (("reference", CAM_NM_REFERENCE), ("orthogonal", CAM_NM_ORTHOGONAL)):
            ks = []
            for sd in range(args.kappa_seeds):
                r = h_h2_kappa(args.seed + sd, mode, args.scale, args.kelvin)
                ks.append(r["kappa"])
                print(json.dumps(dict(nm=name, seed=args.seed + sd, kappa=r["kappa"], errors=r["errors"])), flush=True)
            ks = np.array(ks)
            out[name] = dict(kappa_mean=float(ks.mean()), kappa_sem=float(ks.std(ddof=1) / np.sqrt(len(ks))),
                             kappas=[float(k) for k in ks])
        print(json.dumps(out))
        raise SystemExit(0)
    out = h_h2_example(args.scale, args.kelvin, args.seed, args.pmf_method)
    res = dict(kelvin=args.kelvin, scale=args.scale, pmf_method=args.pmf_method, xi_barrier=out["xi_barrier"],
               dW_kcal_mol=(out["pmf_max"] - out["pmf_min"]) * 627.5094743, kappa=out["kappa"],
               k_cm3_molecule_s=out["k_cm3_molecule_s"], errors=out["errors"], wall_s=out["wall_s"])
    if "pmf_wham" in out:
        w = out["pmf_wham"]
        res["wham"] = dict(iterations=out["wham_iterations"], xi_barrier=float(out["wham_coord"][int(np.argmax(w))]),
                           max_kcal_mol=float(w.max() * 627.5094743))
    print(json.dumps(res))
