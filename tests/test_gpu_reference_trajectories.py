"""The GPU engine against trajectories of the reference's own step.

The fixtures hold frames of src/verlet.f90 compiled from the reference's
sources (tests/golden/make_reference_trajectories.py): C1 (H + H2, 16 beads,
NVE with transrot), a CH4 + H recrossing child (32 beads, xi every step) and
OH + H2 (8 beads).  The engine starts from the same inputs; its positions
must stay within 1e-10 of the reference's scale over the 100 steps, for each
normal-mode kernel that applies."""
import dataclasses
import os

import numpy as np
import pytest

import caracal_amd as cam
from caracal_amd import workloads

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _frames(e, steps):
    out, done = [], 0
    for k in steps:
        e.run(int(k) - done)
        done = int(k)
        out.append(e.get_state())
    return np.array([o[0][0] for o in out]), np.array([o[1][0] for o in out])


def _check(name, q, p, f):
    eq = np.abs(q - f["q"]).max() / np.abs(f["q"]).max()
    ep = np.abs(p - f["p"]).max() / np.abs(f["p"]).max()
    print(f"{name}: max rel position error {eq:.2e}, momentum {ep:.2e} over steps {list(f['steps'])}", flush=True)
    assert eq <= 1e-10 and ep <= 1e-9


@pytest.mark.parametrize("nmk", [cam.CAM_NMK_VALU, cam.CAM_NMK_MFMA])
def test_c1_h3_nve_matches_reference(nmk):
    f = dict(np.load(os.path.join(GOLDEN, "c1_h3_nve_reference.npz")))
    nb = f["q0"].shape[0]
    with cam.RPMDEngine(cam.CAM_SURFACE_H3, nb, 1, float(f["dt"]), f["masses"], remove_com=True,
                        nm_kernel=nmk) as e:
        e.set_state(f["q0"][None], f["p0"][None], [float(f["beta"])])
        q, p = _frames(e, f["steps"])
    _check(f"C1 nm_kernel={nmk}", q, p, f)


@pytest.mark.parametrize("nmk", [cam.CAM_NMK_VALU, cam.CAM_NMK_MFMA, cam.CAM_NMK_FOLDED])
def test_ch4h_child_matches_reference(nmk):
    f = dict(np.load(os.path.join(GOLDEN, "ch4h_child_reference.npz")))
    nb = f["q0"].shape[0]
    with cam.RPMDEngine(cam.CAM_SURFACE_CH4H, nb, 1, float(f["dt"]), f["masses"], ensemble=cam.CAM_ENS_CHILD,
                        reaction=workloads.ch4h_reaction(), nm_kernel=nmk) as e:
        e.set_state(f["q0"][None], f["p0"][None], [float(f["beta"])])
        e.begin_children()
        q, p = _frames(e, f["steps"])
        xi = e.observables()["xi"][0]
    _check(f"CH4H child nm_kernel={nmk}", q, p, f)
    assert abs(xi - f["xi"][-1]) < 1e-10


@pytest.mark.parametrize("prec", [cam.CAM_PREC_FP64, cam.CAM_PREC_MIXED])
def test_h3_nve_6_beads_matches_reference(prec):
    """A bead count that is not a power of two (the reference's FFTW
    transforms take any): 6 beads, NVE with transrot; mixed precision to its
    binary32 corrections' level."""
    f = dict(np.load(os.path.join(GOLDEN, "h3_nve_6beads_reference.npz")))
    nb = f["q0"].shape[0]
    assert nb == 6
    with cam.RPMDEngine(cam.CAM_SURFACE_H3, nb, 1, float(f["dt"]), f["masses"], remove_com=True,
                        precision=prec) as e:
        e.set_state(f["q0"][None], f["p0"][None], [float(f["beta"])])
        q, p = _frames(e, f["steps"])
    if prec == cam.CAM_PREC_FP64:
        _check("H3 6 beads", q, p, f)
    else:
        assert np.abs(q - f["q"]).max() <= 1e-6 * np.abs(f["q"]).max()


@pytest.mark.parametrize("nmk", [cam.CAM_NMK_VALU, cam.CAM_NMK_FOLDED])
def test_ch4h_child_12_beads_matches_reference(nmk):
    f = dict(np.load(os.path.join(GOLDEN, "ch4h_child_12beads_reference.npz")))
    nb = f["q0"].shape[0]
    assert nb == 12
    with cam.RPMDEngine(cam.CAM_SURFACE_CH4H, nb, 1, float(f["dt"]), f["masses"], ensemble=cam.CAM_ENS_CHILD,
                        reaction=workloads.ch4h_reaction(), nm_kernel=nmk) as e:
        e.set_state(f["q0"][None], f["p0"][None], [float(f["beta"])])
        e.begin_children()
        q, p = _frames(e, f["steps"])
        xi = e.observables()["xi"][0]
    _check(f"CH4H child 12 beads nm_kernel={nmk}", q, p, f)
    assert abs(xi - f["xi"][-1]) < 1e-10


def test_oh3_nve_matches_reference():
    f = dict(np.load(os.path.join(GOLDEN, "oh3_nve_reference.npz")))
    nb = f["q0"].shape[0]
    with cam.RPMDEngine(cam.CAM_SURFACE_OH3, nb, 1, float(f["dt"]), f["masses"], remove_com=True) as e:
        e.set_state(f["q0"][None], f["p0"][None], [float(f["beta"])])
        q, p = _frames(e, f["steps"])
    _check("OH3", q, p, f)


def test_brh2_nve_matches_reference():
    """H + HBr at the BrH2 saddle, 8 beads, NVE with transrot: the engine's
    Jacobi eigen-solve against the reference's EISPACK one, step by step."""
    f = dict(np.load(os.path.join(GOLDEN, "brh2_nve_reference.npz")))
    nb = f["q0"].shape[0]
    with cam.RPMDEngine(cam.CAM_SURFACE_BRH2, nb, 1, float(f["dt"]), f["masses"], remove_com=True) as e:
        e.set_state(f["q0"][None], f["p0"][None], [float(f["beta"])])
        q, p = _frames(e, f["steps"])
    _check("BrH2", q, p, f)


def test_o3_nve_matches_reference():
    """O + O2 at the exchange saddle of the O3 surface, 8 beads, NVE with
    transrot."""
    f = dict(np.load(os.path.join(GOLDEN, "o3_nve_reference.npz")))
    nb = f["q0"].shape[0]
    with cam.RPMDEngine(cam.CAM_SURFACE_O3, nb, 1, float(f["dt"]), f["masses"], remove_com=True) as e:
        e.set_state(f["q0"][None], f["p0"][None], [float(f["beta"])])
        q, p = _frames(e, f["steps"])
    _check("O3", q, p, f)


@pytest.mark.parametrize("schedule", [cam.CAM_SCHED_FUSED, cam.CAM_SCHED_SPLIT])
def test_h3_umbrella_window_matches_reference(schedule):
    """An umbrella window as calc_rate runs it (verlet constrain = 0): mdinit's
    biased start forces, the bias with its metric term every step, and
    transrot after every step."""
    f = dict(np.load(os.path.join(GOLDEN, "h3_window_reference.npz")))
    nb = f["q0"].shape[0]
    rx = dataclasses.replace(workloads.h3_reaction(), xi_ideal=float(f["xi_ideal"]))
    with cam.RPMDEngine(cam.CAM_SURFACE_H3, nb, 1, float(f["dt"]), f["masses"], ensemble=cam.CAM_ENS_UMBRELLA,
                        reaction=rx, remove_com=True, schedule=schedule) as e:
        e.set_state(f["q0"][None], f["p0"][None], [float(f["beta"])])
        e.set_windows([float(f["xi_ideal"])], [float(f["k_force"])])
        e.begin_umbrella(bias=True)
        q, p = _frames(e, f["steps"])
        avg, _ = e.umbrella_stats()
    _check(f"H3 window schedule={schedule}", q, p, f)
    assert abs(avg[0] - f["xi"].mean()) < 1e-10


def test_h3_nose_hoover_chain_matches_reference():
    f = dict(np.load(os.path.join(GOLDEN, "h3_nhc_reference.npz")))
    nb = f["q0"].shape[0]
    with cam.RPMDEngine(cam.CAM_SURFACE_H3, nb, 1, float(f["dt"]), f["masses"], ensemble=cam.CAM_ENS_NHC,
                        remove_com=True, nose_q=float(f["nose_q"])) as e:
        e.set_state(f["q0"][None], f["p0"][None], [float(f["beta"])])
        q, p = _frames(e, f["steps"])
    _check("H3 Nose-Hoover chain", q, p, f)


def test_oh3_child_at_xi_half_matches_reference():
    """Dividing surface at xi_ideal = 0.5, so the child's xi mixes s1 with s0
    of the mass-weighted fragment centres."""
    f = dict(np.load(os.path.join(GOLDEN, "oh3_child_xi05_reference.npz")))
    nb = f["q0"].shape[0]
    rx = dataclasses.replace(workloads.oh3_reaction(), xi_ideal=float(f["xi_ideal"]))
    with cam.RPMDEngine(cam.CAM_SURFACE_OH3, nb, 1, float(f["dt"]), f["masses"], ensemble=cam.CAM_ENS_CHILD,
                        reaction=rx) as e:
        e.set_state(f["q0"][None], f["p0"][None], [float(f["beta"])])
        e.begin_children()
        q, p = _frames(e, f["steps"])
        xi = e.observables()["xi"][0]
    _check("OH3 child xi_ideal 0.5", q, p, f)
    assert abs(xi - f["xi"][-1]) < 1e-10
