"""GPU parity of the reference's QMDFF (cam_qmdff_create, ff_eg + ff_nonb +
ff_hb on the device) and of the box engine running it.

* The molecules and the 16-molecule periodic box of the fixtures against the
  reference's own energies and gradients (written by the compiled reference).
* The example's 1125-atom ethanol box (every molecule with the ethanol
  molecule's parameters) against the oracle, FP64 and mixed precision.
* Ring-polymer steps of that box with the Nose-Hoover chain and the
  reference's periodic wrap against the oracle's."""
import numpy as np
import pytest
import torch

import caracal_amd as cam
import oracle_lib as ol
import qmdff_fixtures as qf
from box_oracle import BoxRPMD, QmdffFF
from caracal_amd import ff as cff
from caracal_amd.units import mass_au

pytestmark = pytest.mark.gpu

AMU = {1: 1.008, 6: 12.011, 8: 15.999, 7: 14.007, 16: 32.06}


def _forces(ff, xs, precision=cam.CAM_PREC_FP64):
    x = torch.tensor(np.ascontiguousarray(xs), device="cuda")
    e, g, _, _ = ff.forces(None, x, precision=precision)
    torch.cuda.synchronize()
    return e.cpu().numpy(), g.cpu().numpy()


@pytest.mark.parametrize("name", qf.SYSTEMS)
def test_qmdff_matches_reference_fixture(name):
    p, xs, ref = qf.load(name)
    print(f"{name}: {p['natoms']} atoms, {len(p['bond'])} bonds, {len(p['tors'])} torsions, "
          f"{len(p['hb'])} listed hydrogen bonds, {len(p['nci'])} nci pairs, nmols {p['nmols']}", flush=True)
    ff = cff.ForceField.qmdff(p)
    e, g = _forces(ff, xs)
    ff.close()
    for k in range(len(xs)):
        de = abs(e[k] - ref["e"][k])
        dg = np.abs(g[k] - ref["g"][k]).max() / np.abs(ref["g"][k]).max()
        print(f"  structure {k}: E {e[k]:.12f} (reference {ref['e'][k]:.12f}) |dE| {de:.1e}, rel dg {dg:.1e}")
        assert de <= 1e-11 * max(1.0, abs(ref["e"][k]))
        assert dg <= 1e-11


def _box(nb, seed=3, spread=0.05):
    p, x = qf.ethanol_box()
    rng = np.random.default_rng(seed)
    xs = x[None] + rng.normal(0.0, spread, (nb,) + x.shape)
    return p, xs


def test_ethanol_box_matches_oracle_fp64_and_mixed():
    p, xs = _box(3)
    print(f"ethanol box: {p['natoms']} atoms, 3 beads", flush=True)
    ff = cff.ForceField.qmdff(p)
    e, g = _forces(ff, xs)
    em, gm = _forces(ff, xs, cam.CAM_PREC_MIXED)
    e2, g2 = _forces(ff, xs)
    ff.close()
    assert np.array_equal(e, e2) and np.array_equal(g, g2)          # bitwise reproducible
    o = ol.QmdffOracle(p)
    for b in range(len(xs)):
        er, gr, _ = o.eval(xs[b])
        dg = np.abs(g[b] - gr).max() / np.abs(gr).max()
        dgm = np.abs(gm[b] - gr).max() / np.abs(gr).max()
        print(f"  bead {b}: E {e[b]:.10f} (oracle {er:.10f}), rel dg {dg:.1e}; mixed: dE {abs(em[b] - er):.1e}, "
              f"rel dg {dgm:.1e}")
        assert abs(e[b] - er) <= 1e-11 * abs(er)
        assert dg <= 1e-11
        assert dgm <= 1e-5
        assert abs(em[b] - er) <= 1e-6 * abs(er)


def test_ethanol_box_ring_polymer_nhc_matches_oracle():
    nb, steps = 4, 6
    p, xs = _box(nb, seed=5, spread=0.02)
    m = mass_au([AMU[int(z)] for z in p["at"]])
    beta = float(cam.beta_of(200.0))
    dt = 0.5 * cam.FS_TO_AU
    rng = np.random.default_rng(6)
    pm = rng.normal(size=xs.shape) * np.sqrt(m[None, :, None] * nb / beta)
    ff = cff.ForceField.qmdff(p)
    eng = cff.BoxEngine(ff, nb, dt, m, beta, None, precision=cam.CAM_PREC_FP64, nm_mode=cam.CAM_NM_REFERENCE)
    eng.set_nhc(20 * dt)
    eng.set_state(xs, pm)
    eng.run(steps)
    q, pq = eng.get_state()
    eng.close()
    ff.close()
    ref = BoxRPMD(QmdffFF(ol.QmdffOracle(p)), np.zeros(10), xs, pm, dt, m, beta, nose_q=20 * dt,
                  wrap_box=p["box"])
    for _ in range(steps):
        ref.step()
    err = np.abs(q - ref.q).max() / np.abs(ref.q).max()
    perr = np.abs(pq - ref.p).max() / np.abs(ref.p).max()
    print(f"ethanol box, {nb} beads, {steps} NHC steps: rel dq {err:.1e}, dp {perr:.1e}")
    assert err <= 1e-11 and perr <= 1e-9


def test_ethanol16_box_dynamics_match_reference():
    """The box engine with the Nose-Hoover chain, the periodic wrap and transrot
    against 20 steps of the reference's own dynamic.x verlet on 16 molecules
    of the ethanol box (4 beads, real parameters, fixture from the compiled
    reference)."""
    import os
    f = dict(np.load(os.path.join(qf.GOLDEN, "qmdff_ethanol16_nhc_reference.npz")))
    p, _, _ = qf.load("ethanol16")
    nb = f["q0"].shape[0]
    ff = cff.ForceField.qmdff(p)
    eng = cff.BoxEngine(ff, nb, float(f["dt"]), f["masses"], float(f["beta"]), None, precision=cam.CAM_PREC_FP64,
                        nm_mode=cam.CAM_NM_REFERENCE)
    eng.set_nhc(float(f["nose_q"]))
    eng.set_transrot(True)
    eng.set_state(f["q0"], f["p0"])
    done = 0
    for k, step in enumerate(f["steps"]):
        eng.run(int(step) - done)
        done = int(step)
        q, pq = eng.get_state()
        eq = np.abs(q - f["q"][k]).max() / np.abs(f["q"][k]).max()
        ep = np.abs(pq - f["p"][k]).max() / np.abs(f["p"][k]).max()
        print(f"step {step}: rel dq {eq:.1e}, dp {ep:.1e}", flush=True)
        assert eq <= 1e-10 and ep <= 1e-9
    ep_beads, _, _ = eng.energies()
    assert abs(ep_beads.sum() - f["epot"][-1]) <= 1e-10 * max(1.0, abs(f["epot"][-1]))
    eng.close()
    ff.close()
