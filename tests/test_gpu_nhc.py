"""GPU parity of the Nose-Hoover chain ensemble (dynamic.x with thermostat 2,
src/nhc.f90): the chain before the first and after the second half kick of
every step, against the oracle, over launches that split the run at
arbitrary steps; set_state restarts the chains as mdinit does."""
import numpy as np
import pytest

import caracal_amd as cam
import oracle_lib as ol
from caracal_amd import workloads

pytestmark = pytest.mark.gpu


def _batch(surface, nb, ntraj, seed):
    if surface == cam.CAM_SURFACE_H3:
        w = workloads.h3_nve(nbeads=nb)
        rng = np.random.default_rng(seed)
        q = np.stack([w["q"]] * ntraj) + rng.normal(0, 0.01, (ntraj,) + w["q"].shape)
    else:
        w = workloads.ch4h_children(ntraj=ntraj, nbeads=nb, seed=seed)
        q = w["q"]
        rng = np.random.default_rng(seed)
    beta = np.full(ntraj, float(np.atleast_1d(w["beta"])[0])) * rng.uniform(0.6, 1.2, ntraj)
    # momenta 1.5 times too hot, so that the chains have work to do
    p = 1.5 * rng.normal(size=q.shape) * np.sqrt(w["masses"][None, None, :, None] * nb / beta[:, None, None, None])
    return w, q, p, beta


@pytest.mark.parametrize("surface,nb", [(cam.CAM_SURFACE_H3, 8), (cam.CAM_SURFACE_CH4H, 4),
                                        (cam.CAM_SURFACE_CH4H, 16), (cam.CAM_SURFACE_H3, 1),
                                        (cam.CAM_SURFACE_H3, 6), (cam.CAM_SURFACE_CH4H, 24)])
def test_nhc_matches_oracle(surface, nb):
    # 6 and 24 beads leave slots of a block past its last ring polymer: enough
    # trajectories that a block is full and the next one holds the rest
    ntraj, nsteps = (50 if nb == 6 else 5 if nb != 24 else 5), 50
    w, q, p, beta = _batch(surface, nb, ntraj, seed=nb + 3)
    nose_q = 20.0 * w["dt"]
    e = cam.RPMDEngine(surface, nb, ntraj, w["dt"], w["masses"], ensemble=cam.CAM_ENS_NHC, nose_q=nose_q,
                       steps_per_launch=7)
    e.set_state(q, p, beta)
    e.run(nsteps)
    qe, pe = e.get_state()
    s = ol.make_sys(surface, nb, w["dt"], w["masses"])
    for t in range(ntraj):
        tr = ol.OracleTraj(s, q[t], p[t], float(beta[t]), ensemble=5)
        tr.nhc_begin(nose_q)
        for _ in range(nsteps):
            tr.step()
        err = np.abs(qe[t] - tr.q).max() / np.abs(tr.q).max()
        perr = np.abs(pe[t] - tr.p).max() / np.abs(tr.p).max()
        print(f"surface={surface} nb={nb} traj={t}: rel dq {err:.1e} dp {perr:.1e}; chain v1 {tr.vnh[0]:.3e}")
        assert err <= 1e-10 and perr <= 1e-9
        assert abs(tr.vnh[0]) > 0                     # the chain moved
    # a new state restarts the chains: the same run again gives the same trajectories
    e.set_state(q, p, beta)
    e.run(nsteps)
    q2, p2 = e.get_state()
    assert np.array_equal(q2, qe) and np.array_equal(p2, pe)
    e.close()


def test_nhc_autotune_and_rejections():
    w, q, p, beta = _batch(cam.CAM_SURFACE_H3, 8, 40, seed=1)
    e = cam.RPMDEngine(cam.CAM_SURFACE_H3, 8, 40, w["dt"], w["masses"], ensemble=cam.CAM_ENS_NHC,
                       nose_q=20.0 * w["dt"])
    e.set_state(q, p, beta)
    q0, p0 = e.get_state()
    trials, chosen = e.autotune(10)
    assert all(t["split"] == 0 and t["nm_kernel"] == cam.CAM_NMK_VALU for t in trials)
    q1, p1 = e.get_state()
    assert np.array_equal(q0, q1) and np.array_equal(p0, p1)
    e.close()
    for kw, msg in ((dict(nose_q=0.0), "nose_q"), (dict(nose_q=1.0, schedule=cam.CAM_SCHED_SPLIT), "fused"),
                    (dict(nose_q=1.0, nm_kernel=cam.CAM_NMK_MFMA), "fused")):
        with pytest.raises(cam.CamError, match=msg):
            cam.RPMDEngine(cam.CAM_SURFACE_H3, 16, 1, w["dt"], w["masses"], ensemble=cam.CAM_ENS_NHC, **kw)
