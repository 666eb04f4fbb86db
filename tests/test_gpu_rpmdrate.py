"""The RPMD rate constant end to end on the GPU (caracal_amd.rpmdrate), on
H + H2 and on OH + H2 (the second test below).  First the
H + H2 calculation of examples/calc_rate/h+h2/rate.key at 300 K and 8 beads,
with every phase shortened and the windows spaced 0.05 (and extended to
xi = 1.15, so that the barrier has windows on both sides).  Checks what the
physics fixes: a PMF barrier at the symmetric transition state (xi near 1)
whose height is the surface's 9.6 kcal/mol classical barrier plus the
translational free energy lost between R_inf and the transition state (a few
kcal/mol, less the zero-point effects), window averages that follow their
windows, a recrossing factor in (0, 1], and a positive, finite rate."""
import time

import numpy as np
import pytest

import caracal_amd as cam
from caracal_amd import workloads
from caracal_amd.rpmdrate import PmfKeys, RecrossKeys, RPMDRate, UmbrellaKeys
from caracal_amd.units import FS_TO_AU, mass_au

pytestmark = pytest.mark.gpu


def test_h_plus_h2_rate_end_to_end():
    ts = workloads.H3_TS
    ts_energy = float(cam.egrad_h3(ts[None])[0][0])
    r = RPMDRate(cam.CAM_SURFACE_H3, 8, 0.1 * FS_TO_AU, mass_au([1.008] * 3), workloads.h3_reaction(), 300.0,
                 umbrella=UmbrellaKeys(bias=0.05, bonds=(-0.05, 1.15), dist=0.05, gen_steps=300, equi_steps=300,
                                       sample_steps=2000, sample_trajs=2),
                 pmf=PmfKeys(xi_range=(-0.05, 1.15), bins=601),
                 recross=RecrossKeys(equi_steps=500, child_total=256, child_interval=100, child_perpoint=64,
                                     child_steps=300, nparents=4),
                 andersen_step=80, n_paths=2, seed=7)
    t0 = time.perf_counter()
    out = r.run(ts, ts_energy)
    wall = time.perf_counter() - t0
    barrier = (out["pmf_max"] - out["pmf_min"]) * 627.5094743
    print(f"wall {wall:.1f} s; xi_barrier {out['xi_barrier']:.3f}; dW {barrier:.2f} kcal/mol; "
          f"kappa {out['kappa']:.3f}; k(T) {out['k_cm3_molecule_s']:.3e} cm^3/(molecule s); "
          f"failed segments {out['errors']}")
    assert 0.95 <= out["xi_barrier"] <= 1.06
    assert 8.0 <= barrier <= 16.0
    # the biased averages sit between their window centre and the barrier top
    assert np.abs(out["average"] - out["xi_wins"]).max() < 0.1
    assert 0.2 < out["kappa"] <= 1.05
    assert np.isfinite(out["k_cm3_molecule_s"]) and out["k_cm3_molecule_s"] > 0


def test_oh_plus_h2_rate_end_to_end():
    """The same pipeline on the OH + H2 surface (src/egrad_oh3.f) at 300 K:
    the PMF barrier sits near xi = 1 and is the surface's 6.09 kcal/mol
    classical barrier plus the translational free energy lost between R_inf
    and the saddle, less zero-point effects; two equivalent abstraction
    paths; kappa in (0, 1]; a positive, finite rate."""
    ts = workloads.OH3_TS
    ts_energy = float(cam.egrad_oh3(ts[None])[0][0])
    r = RPMDRate(cam.CAM_SURFACE_OH3, 8, 0.1 * FS_TO_AU, mass_au(workloads.MASSES_OH3_AMU), workloads.oh3_reaction(),
                 300.0,
                 umbrella=UmbrellaKeys(bias=0.05, bonds=(-0.05, 1.15), dist=0.05, gen_steps=300, equi_steps=300,
                                       sample_steps=2000, sample_trajs=2),
                 pmf=PmfKeys(xi_range=(-0.05, 1.15), bins=601),
                 recross=RecrossKeys(equi_steps=500, child_total=256, child_interval=100, child_perpoint=64,
                                     child_steps=300, nparents=4),
                 andersen_step=80, n_paths=2, seed=3)
    t0 = time.perf_counter()
    out = r.run(ts, ts_energy)
    wall = time.perf_counter() - t0
    barrier = (out["pmf_max"] - out["pmf_min"]) * 627.5094743
    print(f"OH + H2: wall {wall:.1f} s; xi_barrier {out['xi_barrier']:.3f}; dW {barrier:.2f} kcal/mol; "
          f"kappa {out['kappa']:.3f}; k(T) {out['k_cm3_molecule_s']:.3e} cm^3/(molecule s); "
          f"failed segments {out['errors']}")
    assert 0.9 <= out["xi_barrier"] <= 1.1
    assert 5.0 <= barrier <= 15.0
    assert np.abs(out["average"] - out["xi_wins"]).max() < 0.1
    assert 0.2 < out["kappa"] <= 1.05
    assert np.isfinite(out["k_cm3_molecule_s"]) and out["k_cm3_molecule_s"] > 0


def _rate_run(surface, ts, ts_energy, masses, reaction, n_paths):
    r = RPMDRate(surface, 8, 0.1 * FS_TO_AU, mass_au(masses), reaction, 300.0,
                 umbrella=UmbrellaKeys(bias=0.05, bonds=(-0.05, 1.15), dist=0.05, gen_steps=300, equi_steps=300,
                                       sample_steps=2000, sample_trajs=2),
                 pmf=PmfKeys(xi_range=(-0.05, 1.15), bins=601),
                 recross=RecrossKeys(equi_steps=500, child_total=256, child_interval=100, child_perpoint=64,
                                     child_steps=300, nparents=4),
                 andersen_step=80, n_paths=n_paths, seed=3)
    t0 = time.perf_counter()
    out = r.run(ts, ts_energy)
    return out, time.perf_counter() - t0


def test_h_plus_hbr_rate_end_to_end():
    """The pipeline on the BrH2 DIM-3C surface (src/egrad_brh2.f) for
    H + HBr -> H2 + Br at 300 K: an early, low barrier (1.82 kcal/mol above
    H + HBr at the collinear saddle) plus the translational free energy lost
    between R_inf and the saddle; one abstraction path; kappa in (0, 1]; a
    positive, finite rate."""
    ts = workloads.BRH2_TS
    out, wall = _rate_run(cam.CAM_SURFACE_BRH2, ts, float(cam.egrad_brh2(ts[None])[0][0]),
                          workloads.MASSES_BRH2_AMU, workloads.brh2_reaction(), 1)
    barrier = (out["pmf_max"] - out["pmf_min"]) * 627.5094743
    print(f"H + HBr: wall {wall:.1f} s; xi_barrier {out['xi_barrier']:.3f}; dW {barrier:.2f} kcal/mol; "
          f"kappa {out['kappa']:.3f}; k(T) {out['k_cm3_molecule_s']:.3e} cm^3/(molecule s); "
          f"failed segments {out['errors']}")
    assert 0.8 <= out["xi_barrier"] <= 1.1
    assert 0.5 <= barrier <= 10.0
    assert np.abs(out["average"] - out["xi_wins"]).max() < 0.1
    assert 0.1 < out["kappa"] <= 1.05
    assert np.isfinite(out["k_cm3_molecule_s"]) and out["k_cm3_molecule_s"] > 0


