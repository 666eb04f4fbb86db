"""GPU parity of the recrossing parent (ensemble PARENT, verlet with
constrain = 1): the engine's start (mdinit with the umbrella bias) and its
constrained steps with Andersen draws, against the oracle trajectory by
trajectory."""
import numpy as np
import pytest

import caracal_amd as cam
import oracle_lib as ol
from caracal_amd import workloads

pytestmark = pytest.mark.gpu


def _parents(surface, nb, ntraj, seed):
    if surface == cam.CAM_SURFACE_H3:
        w = workloads.h3_nve(nbeads=nb)
        rng = np.random.default_rng(seed)
        q = np.stack([w["q"]] * ntraj) + rng.normal(0, 0.01, (ntraj,) + w["q"].shape)
        beta = np.full(ntraj, w["beta"]) * rng.uniform(0.7, 1.4, ntraj)
        rx = workloads.h3_reaction()
    elif surface == cam.CAM_SURFACE_OH3:
        w = workloads.oh3_nve(nbeads=nb)
        rng = np.random.default_rng(seed)
        q = np.stack([w["q"]] * ntraj) + rng.normal(0, 0.01, (ntraj,) + w["q"].shape)
        beta = np.full(ntraj, w["beta"]) * rng.uniform(0.7, 1.4, ntraj)
        rx = workloads.oh3_reaction()
    else:
        w = workloads.ch4h_children(ntraj=ntraj, nbeads=nb, seed=seed)
        q, beta, rx = w["q"], w["beta"], w["reaction"]
        rng = np.random.default_rng(seed)
    p = rng.normal(size=q.shape) * np.sqrt(w["masses"][None, None, :, None] * nb / beta[:, None, None, None])
    return w, q, p, beta, rx


@pytest.mark.parametrize("surface,nb,nmk", [(cam.CAM_SURFACE_CH4H, 8, cam.CAM_NMK_VALU),
                                            (cam.CAM_SURFACE_CH4H, 32, cam.CAM_NMK_MFMA),
                                            (cam.CAM_SURFACE_H3, 16, cam.CAM_NMK_VALU),
                                            (cam.CAM_SURFACE_OH3, 8, cam.CAM_NMK_VALU)])
def test_parent_matches_oracle(surface, nb, nmk):
    ntraj, andersen_step, k_force, seed = 5, 7, 0.37, 11
    w, q, p, beta, rx = _parents(surface, nb, ntraj, seed=nb)
    e = cam.RPMDEngine(surface, nb, ntraj, w["dt"], w["masses"], ensemble=cam.CAM_ENS_PARENT, reaction=rx,
                       andersen_step=andersen_step, seed=seed, nm_kernel=nmk)
    e.set_state(q, p, beta)
    e.begin_parent(k_force)
    s = ol.make_sys(surface, nb, w["dt"], w["masses"])
    orx = ol.make_rx(rx)
    refs = []
    for t in range(ntraj):
        tr = ol.OracleTraj(s, q[t], p[t], float(beta[t]), rx=orx, ensemble=3, andersen_step=andersen_step,
                           seed=seed, traj=t)
        tr.parent_begin(k_force)
        refs.append(tr)
    _, g0 = e.get_forces()
    for t in range(ntraj):
        np.testing.assert_allclose(g0[t], refs[t].g, rtol=0, atol=1e-12 * np.abs(refs[t].g).max())
    for block in range(3):
        e.run(15)
        statuses = [[tr.step() for _ in range(15)] for tr in refs]
        qe, pe = e.get_state()
        obs = e.observables()
        for t, tr in enumerate(refs):
            assert not any(statuses[t])
            err = np.abs(qe[t] - tr.q).max() / np.abs(tr.q).max()
            perr = np.abs(pe[t] - tr.p).max() / np.abs(tr.p).max()
            print(f"surface={surface} nb={nb} traj={t} step={15 * (block + 1)}: rel dq {err:.1e} dp {perr:.1e}")
            assert err <= 1e-10 and perr <= 1e-9
        # on the dividing surface (constrain_q stops once |xi| < 1e-10), no failed constraint
        assert np.abs(obs["xi"]).max() < 1.5e-10
        assert not (obs["flags"] & 2).any()
    e.close()


def test_parent_andersen_step_and_projection():
    """With andersen_step 5, steps 5 and 10 end on the counter-based draw;
    every other step ends with no momentum along the constraint
    direction (constrain_p)."""
    w, q, p, beta, rx = _parents(cam.CAM_SURFACE_CH4H, 8, 3, seed=2)
    e = cam.RPMDEngine(cam.CAM_SURFACE_CH4H, 8, 3, w["dt"], w["masses"], ensemble=cam.CAM_ENS_PARENT, reaction=rx,
                       andersen_step=5, seed=9)
    e.set_state(q, p, beta)
    e.begin_parent()
    s = ol.make_sys(2, 8, w["dt"], w["masses"])
    orx = ol.make_rx(rx)
    refs = []
    for t in range(3):
        tr = ol.OracleTraj(s, q[t], p[t], float(beta[t]), rx=orx, ensemble=3, andersen_step=5, seed=9, traj=t)
        tr.parent_begin()
        refs.append(tr)
    for step in range(1, 11):
        e.run(1)
        qe, pe = e.get_state()
        for t in range(3):
            refs[t].step()
            if step % 5 == 0:
                ref = np.array([[[ol.andersen_normal(9, t, step, 3 * a + d, k) for d in range(3)] for a in range(6)]
                                for k in range(8)])
                ref *= np.sqrt(w["masses"][None, :, None] / (beta[t] / 8))
                # the same counter-based draw; the device's log/sincos may differ from the host's in the last bit
                np.testing.assert_allclose(pe[t], ref, rtol=1e-13, atol=0)
            else:
                # this step's constraint direction: the gradient of xi at the centroid before constrain_q
                dxi = refs[t].dxi
                flux = np.sum(dxi[None] * pe[t] / w["masses"][None, :, None])
                assert abs(flux) < 1e-12 * np.sum(np.abs(dxi[None] * pe[t] / w["masses"][None, :, None]))
    e.close()


def test_parent_autotune_keeps_state_and_split_schedule():
    w, q, p, beta, rx = _parents(cam.CAM_SURFACE_CH4H, 32, 40, seed=3)
    e = cam.RPMDEngine(cam.CAM_SURFACE_CH4H, 32, 40, w["dt"], w["masses"], ensemble=cam.CAM_ENS_PARENT, reaction=rx,
                       andersen_step=10, seed=1)
    e.set_state(q, p, beta)
    with pytest.raises(cam.CamError, match="begin_parent"):
        e.autotune(5)
    with pytest.raises(cam.CamError, match="begin_parent"):
        e.run(1)
    e.begin_parent()
    q0, p0 = e.get_state()
    trials, chosen = e.autotune(5)
    assert {t["nm_kernel"] for t in trials} == {cam.CAM_NMK_VALU, cam.CAM_NMK_MFMA}
    assert all(t["split"] == 1 for t in trials)
    q1, p1 = e.get_state()
    assert np.array_equal(q0, q1) and np.array_equal(p0, p1)
    e.run(20)
    assert np.abs(e.observables()["xi"]).max() < 1.5e-10
    e.close()


def test_parent_rejects_what_it_does_not_cover():
    w = workloads.h3_nve(nbeads=8)
    rx = workloads.h3_reaction()
    for kw, msg in ((dict(schedule=cam.CAM_SCHED_FUSED), "split schedule"), (dict(remove_com=True), "remove_com"),
                    (dict(andersen_step=-1), "andersen_step"), (dict(nm_kernel=cam.CAM_NMK_FOLDED), "FOLDED")):
        with pytest.raises(cam.CamError, match=msg):
            cam.RPMDEngine(cam.CAM_SURFACE_H3, 8, 1, w["dt"], w["masses"], ensemble=cam.CAM_ENS_PARENT, reaction=rx,
                           **kw)
    e = cam.RPMDEngine(cam.CAM_SURFACE_H3, 8, 1, w["dt"], w["masses"])
    e.set_state(w["q"][None], w["p"][None], w["beta"])
    with pytest.raises(cam.CamError, match="PARENT"):
        e.begin_parent()
    e.close()


def test_parent_trajectory_check_arms_after_check_after_steps():
    """The per-step check flags a trajectory whose bead-energy sum exceeds the
    limit, only from step check_after + 1 after the start, and never when
    disarmed; the constraint flag stays clear."""
    w, q, p, beta, rx = _parents(cam.CAM_SURFACE_H3, 8, 6, seed=4)
    e = cam.RPMDEngine(cam.CAM_SURFACE_H3, 8, 6, w["dt"], w["masses"], ensemble=cam.CAM_ENS_PARENT, reaction=rx)
    e.set_state(q, p, beta)
    v, _ = e.get_forces()
    limit = float(np.median(v.sum(axis=1)))       # half of the trajectories lie above it
    e.begin_parent(epot_max=limit, check_after=3)
    e.run(3)
    assert not e.observables()["flags"].any()
    e.run(2)
    vs = e.get_forces()[0].sum(axis=1)
    flags = e.observables()["flags"]
    assert ((flags & 4) != 0).any() and not (flags & 2).any()
    # a trajectory whose last energy is above the limit is flagged
    assert all(flags[t] & 4 for t in range(6) if vs[t] > limit)
    e.begin_parent(epot_max=limit, check_after=-1)
    e.run(5)
    assert not e.observables()["flags"].any()
    e.close()
