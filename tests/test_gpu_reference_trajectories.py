"""The GPU engine against trajectories of the reference's own step.

The fixtures hold frames of src/verlet.f90 compiled from the reference's
sources (tests/golden/make_reference_trajectories.py): C1 (H + H2, 16 beads,
NVE with transrot), a CH4 + H recrossing child (32 beads, xi every step) and
OH + H2 (8 beads).  The engine starts from the same inputs; its positions
must stay within 1e-10 of the reference's scale over the 100 steps, for each
normal-mode kernel that applies."""
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


def test_oh3_nve_matches_reference():
    f = dict(np.load(os.path.join(GOLDEN, "oh3_nve_reference.npz")))
    nb = f["q0"].shape[0]
    with cam.RPMDEngine(cam.CAM_SURFACE_OH3, nb, 1, float(f["dt"]), f["masses"], remove_com=True) as e:
        e.set_state(f["q0"][None], f["p0"][None], [float(f["beta"])])
        q, p = _frames(e, f["steps"])
    _check("OH3", q, p, f)
