"""The oracle's RPMD step against the reference's own step.

Golden trajectories (tests/golden/*_reference.npz) were written by the
reference's src/verlet.f90 -- FFTW normal-mode transforms through MKL,
transrot, umbrella mode 1 / calc_xi, the analytic surfaces -- compiled from
its sources (oracle/Makefile.ref) and driven by oracle/ref_step_harness.f90
(tests/golden/make_reference_trajectories.py).  The oracle must reproduce
them from the same inputs; where the compiled reference is present the
fixtures are also re-derived from it, and longer runs are compared directly.
"""
import dataclasses
import os

import numpy as np
import pytest

import oracle_lib as ol
import refstep
from caracal_amd import workloads

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# fixture, surface, verlet's constrain, thermostat, reaction of the fixture
CASES = [("c1_h3_nve_reference.npz", 1, -1, 0, None),
         ("ch4h_child_reference.npz", 2, 2, 0, workloads.ch4h_reaction),
         ("oh3_nve_reference.npz", 3, -1, 0, None),
         ("h3_window_reference.npz", 1, 0, 0, workloads.h3_reaction),
         ("h3_nhc_reference.npz", 1, -1, 2, None),
         ("oh3_child_xi05_reference.npz", 3, 2, 0, workloads.oh3_reaction),
         ("h3_nve_6beads_reference.npz", 1, -1, 0, None),
         ("ch4h_child_12beads_reference.npz", 2, 2, 0, workloads.ch4h_reaction),
         ("brh2_nve_reference.npz", 4, -1, 0, None),
         ("o3_nve_reference.npz", 5, -1, 0, None)]
SURF = {1: "h3", 2: "ch4h", 3: "oh3", 4: "brh2", 5: "o3"}
# Absolute bound on the energy where the default (1e-12 of its scale) is below
# the rounding of the surface: an O3 bead energy sums polynomial terms of up
# to 1e5 kcal/mol (160 hartree) that cancel to about 0.02 hartree, so either
# evaluation order moves it by ~1e-13; 1e-12 per bead.
EPOT_ATOL = {"o3_nve_reference.npz": 8e-12}


def _fixture(name):
    return dict(np.load(os.path.join(GOLDEN, name)))


def _reaction(f, reaction):
    if reaction is None:
        return None
    return dataclasses.replace(reaction(), xi_ideal=float(f["xi_ideal"]))


def _oracle_frames(f, surface, constrain, thermostat, reaction):
    """The oracle's trajectory from the fixture's inputs: dynamic.x NVE or
    Nose-Hoover chain (transrot after every step), an umbrella window (bias
    start, bias and transrot every step) or a recrossing child."""
    nb = f["q0"].shape[0]
    s = ol.make_sys(surface, nb, float(f["dt"]), f["masses"])
    rx = _reaction(f, reaction)
    beta = float(f["beta"])
    if constrain == 2:
        t = ol.OracleTraj(s, f["q0"], f["p0"], beta, rx=ol.make_rx(rx), ensemble=2)
    elif constrain == 0:
        t = ol.OracleTraj(s, f["q0"], f["p0"], beta, rx=ol.make_rx(rx), ensemble=4, remove_com=True)
        t.umbrella_begin(float(f["k_force"]), bias=True)
    elif thermostat == 2:
        t = ol.OracleTraj(s, f["q0"], f["p0"], beta, ensemble=5, remove_com=True)
        t.nhc_begin(float(f["nose_q"]))
    else:
        t = ol.OracleTraj(s, f["q0"], f["p0"], beta, ensemble=0, remove_com=True)
    qs, ps, xis, eps = [], [], [], []
    for k in range(1, int(f["steps"][-1]) + 1):
        xis.append(t.step())
        eps.append(t.v.sum())
        if k in f["steps"]:
            qs.append(t.q.copy())
            ps.append(t.p.copy())
    return np.array(qs), np.array(ps), np.array(xis), np.array(eps)


@pytest.mark.parametrize("name,surface,constrain,thermostat,reaction", CASES)
def test_oracle_reproduces_reference_trajectory(name, surface, constrain, thermostat, reaction):
    """Positions and momenta within 1e-12 of the reference's scale at steps
    1, 2, 10, 50, 100; the energy every step within 1e-12 of its scale; the
    reaction coordinate of a child (mode 2) or window (mode 1) every step
    within 1e-12 absolute."""
    f = _fixture(name)
    q, p, xi, ep = _oracle_frames(f, surface, constrain, thermostat, reaction)
    eq = np.abs(q - f["q"]).max() / np.abs(f["q"]).max()
    epm = np.abs(p - f["p"]).max() / np.abs(f["p"]).max()
    print(f"{name}: max rel position error {eq:.2e}, momentum {epm:.2e}")
    assert eq < 1e-12 and epm < 1e-12
    assert np.abs(ep - f["epot"]).max() <= EPOT_ATOL.get(name, 1e-12 * np.abs(f["epot"]).max())
    if constrain >= 0:
        np.testing.assert_allclose(xi, f["xi"], rtol=0, atol=1e-12)


def test_transform_matters_to_the_fixture():
    """A plausible mistake in the normal-mode step (the textbook orthogonal
    transform instead of the reference's cosine pair) misses the fixture by
    far more than the tolerance above, so the fixture pins the transform."""
    f = _fixture("c1_h3_nve_reference.npz")
    nb = f["q0"].shape[0]
    # an orthogonal-transform step of the same start, built from the free
    # ring polymer closed form the oracle also exposes
    s = ol.make_sys(1, nb, float(f["dt"]), f["masses"])
    q, p = f["q0"].copy(), f["p0"].copy()
    qs = q.copy()
    ps = p.copy()
    ol.rpmd().oracle_free_rp(ol.ctypes.byref(s), qs.ctypes.data_as(ol._dp), ps.ctypes.data_as(ol._dp),
                             float(f["beta"]))
    t = ol.OracleTraj(s, f["q0"], f["p0"], float(f["beta"]), ensemble=0, remove_com=True)
    t.step()
    # the reference-transform step symmetrises the beads; the unprojected
    # free ring polymer does not
    assert np.abs(t.q - f["q"][0]).max() < 1e-12 * np.abs(f["q"]).max()
    assert np.abs(qs - f["q"][0]).max() > 1e-6


@pytest.mark.skipif(not refstep.available(), reason="compiled reference step not built here")
@pytest.mark.parametrize("name,surface,constrain,thermostat,reaction", CASES)
def test_fixture_is_the_compiled_reference(name, surface, constrain, thermostat, reaction):
    f = _fixture(name)
    kw = {}
    if constrain == 0:
        kw["k_force"] = float(f["k_force"])
    if thermostat == 2:
        kw.update(thermostat=2, kelvin=float(f["kelvin"]), nose_q=float(f["nose_q"]))
    fr, _ = refstep.run(SURF[surface], f["q0"], f["p0"], f["masses"], float(f["dt"]), float(f["beta"]),
                        int(f["steps"][-1]), constrain=constrain, reaction=_reaction(f, reaction), **kw)
    assert np.array_equal(fr["q"][f["steps"] - 1], f["q"])
    assert np.array_equal(fr["epot"], f["epot"])


@pytest.mark.skipif(not refstep.available(), reason="compiled reference step not built here")
def test_oracle_matches_compiled_reference_over_1000_steps():
    """The whole C1 run (1000 steps): positions within 1e-10 relative."""
    h = workloads.h3_nve(nbeads=16)
    fr, _ = refstep.run("h3", h["q"], h["p"], h["masses"], h["dt"], h["beta"], 1000, every=250)
    s = ol.make_sys(1, 16, h["dt"], h["masses"])
    t = ol.OracleTraj(s, h["q"], h["p"], h["beta"], ensemble=0, remove_com=True)
    for k in range(1, 1001):
        t.step()
        if k % 250 == 0:
            err = np.abs(fr["q"][k // 250 - 1] - t.q).max() / np.abs(t.q).max()
            print(f"step {k}: max rel position error {err:.2e}")
            assert err < 1e-10
