"""The oracle's QMDFF (oracle/qmdff.c) against the reference's own ff_eg,
ff_nonb and ff_hb.

The fixtures hold the reference's energies and gradients for the ethanol and
glutathione molecules, the DG-EVB minimum (listed hydrogen bonds) and 16
molecules of the ethanol box in its periodic cell (intermolecular terms and
hydrogen bonds), written by the compiled reference
(tests/golden/make_qmdff_fixtures.py).  Where the compiled reference is
present, the whole 1125-atom box is compared directly, also with Zahn's
Coulomb and without the periodic cell, and the file reader of
caracal_amd.qmdff against the reference's rdsolvff."""
import os

import numpy as np
import pytest

import oracle_lib as ol
import qmdff_fixtures as qf
import refqmdff
from caracal_amd import qmdff

EX = "/root/reference/examples"
BOX = os.path.join(EX, "dynamic/ethanol_box/box.qmdff")


def _check(o, xs, e_ref, g_ref, label, parts_ref=None):
    for k in range(len(xs)):
        e, g, parts = o.eval(xs[k])
        scale = max(1.0, abs(e_ref[k]))
        print(f"{label} structure {k}: E {e:.12f} (reference {e_ref[k]:.12f}), "
              f"max |dg| {np.abs(g - g_ref[k]).max():.2e}")
        assert abs(e - e_ref[k]) <= 1e-12 * scale
        assert np.abs(g - g_ref[k]).max() <= 1e-12 * np.abs(g_ref[k]).max()
        if parts_ref is not None:
            assert abs(parts[0] - parts_ref[0][k]) <= 1e-12 * max(1.0, abs(parts_ref[0][k]))
            assert abs(parts[1] - parts_ref[1][k]) <= 1e-12 * max(1.0, abs(parts_ref[1][k]))


@pytest.mark.parametrize("name", qf.SYSTEMS)
def test_oracle_matches_reference_fixtures(name):
    p, xs, ref = qf.load(name)
    o = ol.QmdffOracle(p)
    _check(o, xs, ref["e"], ref["g"], name, (ref["e_eg"], ref["e_nonb"]))


@pytest.mark.parametrize("name", ("ethanol", "glutathione"))
def test_oracle_gradient_is_the_derivative_of_its_energy(name):
    """Central differences (h = 1e-5 bohr, as gradnum) of the oracle energy on
    the perturbed structure: the analytic gradients of the stretch, bend,
    torsion, out-of-plane, nci and hydrogen-bond terms."""
    p, xs, _ = qf.load(name)
    o = ol.QmdffOracle(p)
    x = xs[1]
    _, g, _ = o.eval(x)
    h = 1e-5
    fd = np.zeros_like(x)
    for i in range(x.shape[0]):
        for c in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, c] += h
            xm[i, c] -= h
            fd[i, c] = (o.eval(xp)[0] - o.eval(xm)[0]) / (2 * h)
    assert np.abs(fd - g).max() <= 1e-6 * np.abs(g).max()


@pytest.mark.skipif(not refqmdff.available(), reason="compiled reference QMDFF not built here")
@pytest.mark.parametrize("variant", ["periodic", "zahn", "cluster"])
def test_oracle_matches_compiled_reference_on_the_ethanol_box(variant):
    t = qmdff.read_qmdff(BOX)
    rng = np.random.default_rng(11)
    xs = np.stack([t["xyz"], t["xyz"] + rng.normal(0, 0.1, t["xyz"].shape)])
    kw = dict(periodic=variant != "cluster", box=27.0, zahn=variant == "zahn")
    e1, e2, e3, g, p = refqmdff.run(BOX, xs, dump=True, **kw)
    _check(ol.QmdffOracle(p), xs, e3, g, f"box {variant}", (e1, e2))


@pytest.mark.skipif(not refqmdff.available(), reason="compiled reference QMDFF not built here")
@pytest.mark.parametrize("rel", ["dynamic/ethanol_box/box.qmdff", "dynamic/glutathione/glutathione.qmdff",
                                 "poly_qmdff/ethanol/ethanol.qmdff", "evbopt/DG-EVB/min1.qmdff",
                                 "dynamic/hydrogen_shift/min1.qmdff"])
def test_reader_equals_reference_rdsolvff(rel):
    path = os.path.join(EX, rel)
    t = qmdff.read_qmdff(path)
    _, _, _, _, p = refqmdff.run(path, t["xyz"][None], dump=True)
    for k in ("at", "q", "molnum", "bond", "vbond", "angl", "vangl", "tors", "vtors", "hb", "vhb", "nci"):
        assert np.array_equal(np.asarray(t[k]), np.asarray(p[k])), k
    assert t["nmols"] == p["nmols"]
    assert np.array_equal(qmdff.scalehb_table(), p["scalehb"])


@pytest.mark.skipif(not os.path.exists(BOX), reason="reference examples not mounted here")
def test_box_file_repeats_the_ethanol_molecule():
    """The example box's terms are the ethanol molecule's, molecule by
    molecule, so the tests' box built from the molecule has the box's bonded
    and nci terms and charges exactly."""
    t = qmdff.read_qmdff(BOX)
    p, x = qf.ethanol_box()
    for k in ("at", "q", "molnum", "bond", "vbond", "angl", "vangl", "tors", "vtors", "nci"):
        assert np.array_equal(np.asarray(t[k]), np.asarray(p[k])), k
    assert np.array_equal(t["xyz"], x)


def test_box_oracle_dynamics_reproduce_the_reference():
    """The box oracle (Nose-Hoover chain, free ring polymer, the periodic wrap,
    QMDFF forces, transrot) against 20 steps of the reference's own dynamic.x
    verlet on the 16-molecule periodic ethanol box with 4 beads."""
    import os
    from box_oracle import BoxRPMD, QmdffFF
    f = dict(np.load(os.path.join(qf.GOLDEN, "qmdff_ethanol16_nhc_reference.npz")))
    p, _, _ = qf.load("ethanol16")
    ref = BoxRPMD(QmdffFF(ol.QmdffOracle(p)), np.zeros(10), f["q0"], f["p0"], float(f["dt"]), f["masses"],
                  float(f["beta"]), nose_q=float(f["nose_q"]), wrap_box=p["box"], remove_com=True)
    k = 0
    for step in range(1, int(f["steps"][-1]) + 1):
        ref.step()
        assert abs(ref.v.sum() - f["epot"][step - 1]) <= 1e-12 * max(1.0, abs(f["epot"][step - 1]))
        if step in f["steps"]:
            eq = np.abs(ref.q - f["q"][k]).max() / np.abs(f["q"][k]).max()
            ep = np.abs(ref.p - f["p"][k]).max() / np.abs(f["p"][k]).max()
            print(f"step {step}: rel dq {eq:.1e}, dp {ep:.1e}")
            assert eq < 1e-12 and ep < 1e-11
            k += 1
