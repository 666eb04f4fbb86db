"""GPU parity of the trajectory-batch engine (C ABI) against the oracle's
verlet, for the reference normal-mode transform, the child trajectories of
the recrossing scheme, Andersen draws and the centroid constraint."""
import numpy as np
import pytest

import caracal_amd as cam
import oracle_lib as ol
from caracal_amd import workloads

pytestmark = pytest.mark.gpu


def _h3_engine(nb, ntraj=1, **kw):
    w = workloads.h3_nve(nbeads=nb)
    e = cam.RPMDEngine(cam.CAM_SURFACE_H3, nb, ntraj, w["dt"], w["masses"], **kw)
    return e, w


@pytest.mark.parametrize("nb,nmk", [(16, cam.CAM_NMK_VALU), (16, cam.CAM_NMK_MFMA), (1, cam.CAM_NMK_AUTO),
                                     (4, cam.CAM_NMK_AUTO), (32, cam.CAM_NMK_MFMA), (64, cam.CAM_NMK_VALU),
                                     (64, cam.CAM_NMK_MFMA)])
def test_h3_nve_trajectory_matches_reference_100_steps(nb, nmk):
    """dynamic.x NVE with transrot, reference normal-mode transform: positions
    within 1e-10 relative of the oracle after 100 steps, with the transform on
    the vector ALUs and on the matrix cores."""
    e, w = _h3_engine(nb, remove_com=True, nm_kernel=nmk)
    e.set_state(w["q"][None], w["p"][None], w["beta"])
    s = ol.make_sys(1, nb, w["dt"], w["masses"])
    t = ol.OracleTraj(s, w["q"], w["p"], w["beta"], ensemble=0, remove_com=True)
    for block in range(4):
        e.run(25)
        for _ in range(25):
            t.step()
        q, p = e.get_state()
        scale = np.abs(t.q).max()
        err = np.abs(q[0] - t.q).max() / scale
        print(f"nb={nb} nm_kernel={nmk} step={25 * (block + 1)} max rel position error {err:.2e}")
        assert err <= 1e-10
        assert np.abs(p[0] - t.p).max() <= 1e-9 * np.abs(t.p).max()
    e.close()


def test_steps_per_launch_and_graph_replay_are_bitwise_invariant():
    w = workloads.ch4h_children(ntraj=24, nbeads=8)
    outs = []
    for spl, runs in ((1, [12]), (5, [12]), (50, [6, 6])):
        e = cam.RPMDEngine(cam.CAM_SURFACE_CH4H, 8, 24, w["dt"], w["masses"], steps_per_launch=spl)
        e.set_state(w["q"], np.random.default_rng(0).normal(size=w["q"].shape) * 20, w["beta"])
        for n in runs:
            e.run(n)
        outs.append(e.get_state()[0])
        e.close()
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


@pytest.mark.parametrize("surface,nb", [(cam.CAM_SURFACE_H3, 4), (cam.CAM_SURFACE_H3, 8), (cam.CAM_SURFACE_H3, 32),
                                        (cam.CAM_SURFACE_CH4H, 4), (cam.CAM_SURFACE_CH4H, 32)])
@pytest.mark.parametrize("prec", [cam.CAM_PREC_FP64, cam.CAM_PREC_MIXED])
def test_fused_steps_equal_single_steps(surface, nb, prec):
    """Steps fused in one launch (state in registers) give bit for bit the
    state of one step per launch, for every surface, bead count and precision,
    on the ragged workload's hot initial conditions (up to 1000 K)."""
    bucket = [b for b in workloads.ragged(ntraj=1200, seed=1) if b[0] == surface and b[1] == nb][0]
    d = bucket[2]
    out = []
    for spl in (1, 10):
        e = cam.RPMDEngine(surface, nb, len(d["beta"]), d["dt"], d["masses"], precision=prec, steps_per_launch=spl)
        e.set_state(d["q"], d["p"], d["beta"])
        e.run(20)
        out.append(e.get_state() + (e.get_forces()[1],))
        e.close()
    for a, b in zip(out[0], out[1]):
        assert np.isfinite(b).all()
        assert np.array_equal(a, b)


def test_orthogonal_mode_equals_reference_mode_on_symmetric_beads():
    nb = 8
    w = workloads.h3_nve(nbeads=nb)
    rev = (-np.arange(nb)) % nb
    q = 0.5 * (w["q"] + w["q"][rev])
    p = 0.5 * (w["p"] + w["p"][rev])
    res = []
    for mode in (cam.CAM_NM_REFERENCE, cam.CAM_NM_ORTHOGONAL):
        e = cam.RPMDEngine(cam.CAM_SURFACE_H3, nb, 1, w["dt"], w["masses"], nm_mode=mode)
        e.set_state(q[None], p[None], w["beta"])
        e.run(50)
        res.append(e.get_state()[0][0])
        e.close()
    assert np.abs(res[0] - res[1]).max() < 1e-10


@pytest.mark.parametrize("nmk", [cam.CAM_NMK_VALU, cam.CAM_NMK_MFMA])
def test_orthogonal_mode_conserves_energy_for_general_beads(nmk):
    """Orthogonal normal modes make the step symplectic for any bead
    configuration: the ring-polymer energy error stays bounded over 400 steps
    and falls as dt^2 (halving dt divides it by ~4)."""
    nb = 16
    w = workloads.h3_nve(nbeads=nb)
    dev = {}
    for div in (1, 2):
        e = cam.RPMDEngine(cam.CAM_SURFACE_H3, nb, 1, w["dt"] / div, w["masses"], nm_mode=cam.CAM_NM_ORTHOGONAL,
                           nm_kernel=nmk)
        e.set_state(w["q"][None], w["p"][None], w["beta"])

        def h():
            o = e.observables()
            return o["epot"][0] + o["ekin"][0] + o["espring"][0]
        h0 = h()
        d = []
        for _ in range(20):
            e.run(20 * div)
            d.append(h() - h0)
        e.close()
        dev[div] = np.abs(np.array(d))
    print(f"nm_kernel={nmk} max |dH| at dt {dev[1].max():.2e}, at dt/2 {dev[2].max():.2e}")
    assert dev[1].max() < 5e-4
    assert 2.5 < dev[1].mean() / dev[2].mean() < 6.0


def test_andersen_draws_match_oracle():
    w = workloads.ch4h_children(ntraj=6, nbeads=4)
    e = cam.RPMDEngine(cam.CAM_SURFACE_CH4H, 4, 6, w["dt"], w["masses"], seed=1234,
                       ensemble=cam.CAM_ENS_CHILD, reaction=w["reaction"])
    e.set_state(w["q"], w["p"], w["beta"])
    e.draw_momenta(draw=77, antithetic_pairs=True)
    _, p = e.get_state()
    s = ol.make_sys(2, 4, w["dt"], w["masses"])
    for t in range(6):
        ref = np.zeros((4, 6, 3))
        ol.rpmd().oracle_andersen(ol.ctypes.byref(s), ref.ctypes.data_as(ol._dp), float(w["beta"][t]), 1234,
                                  t & ~1, 77)
        sign = -1.0 if t & 1 else 1.0
        np.testing.assert_allclose(p[t], sign * ref, rtol=1e-13, atol=1e-12)
    e.close()


def test_constraint_matches_oracle():
    w = workloads.ch4h_children(ntraj=8, nbeads=8)
    e = cam.RPMDEngine(cam.CAM_SURFACE_CH4H, 8, 8, w["dt"], w["masses"], ensemble=cam.CAM_ENS_CHILD,
                       reaction=w["reaction"])
    p0 = np.random.default_rng(2).normal(size=w["q"].shape) * 20
    e.set_state(w["q"], p0, w["beta"])
    st = e.constrain(project_p=True)
    assert not st.any()
    q, p = e.get_state()
    s = ol.make_sys(2, 8, w["dt"], w["masses"])
    rx = ol.make_rx(w["reaction"])
    for t in range(8):
        qt, pt = w["q"][t].copy(), p0[t].copy()
        c = qt.mean(axis=0)
        _, dxi = ol.calc_xi(s, rx, c, 2)
        assert ol.rpmd().oracle_constrain_q(ol.ctypes.byref(s), ol.ctypes.byref(rx), c.ctypes.data_as(ol._dp),
                                            dxi.ctypes.data_as(ol._dp), qt.ctypes.data_as(ol._dp),
                                            pt.ctypes.data_as(ol._dp)) == 0
        ol.rpmd().oracle_constrain_p(ol.ctypes.byref(s), dxi.ctypes.data_as(ol._dp), pt.ctypes.data_as(ol._dp))
        np.testing.assert_allclose(q[t], qt, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(p[t], pt, rtol=1e-10, atol=1e-10)
        assert abs(ol.calc_xi(s, rx, q[t].mean(axis=0), 2)[0]) < 1e-8
    e.close()


def test_children_flags_and_kappa_match_oracle():
    ntraj, nb, nsteps = 16, 8, 60
    w = workloads.ch4h_children(ntraj=ntraj, nbeads=nb)
    e = cam.RPMDEngine(cam.CAM_SURFACE_CH4H, nb, ntraj, w["dt"], w["masses"], ensemble=cam.CAM_ENS_CHILD,
                       reaction=w["reaction"], seed=99, steps_per_launch=25)
    e.set_state(w["q"], w["p"], w["beta"])
    assert not e.constrain(project_p=False).any()
    e.draw_momenta(draw=1, antithetic_pairs=True)
    q0, p0 = e.get_state()
    e.begin_children()
    e.run(nsteps)
    num, den = e.kappa(nsteps)
    q1, _ = e.get_state()
    s = ol.make_sys(2, nb, w["dt"], w["masses"])
    rx = ol.make_rx(w["reaction"])
    num_ref = np.zeros(nsteps)
    den_ref = 0.0
    for t in range(ntraj):
        wt = ol.rpmd().oracle_child_weight(ol.ctypes.byref(s), ol.ctypes.byref(rx),
                                           np.ascontiguousarray(q0[t]).ctypes.data_as(ol._dp),
                                           np.ascontiguousarray(p0[t]).ctypes.data_as(ol._dp), float(w["beta"][t]))
        if wt > 0:
            den_ref += wt
        tr = ol.OracleTraj(s, q0[t], p0[t], float(w["beta"][t]), rx=rx, ensemble=2)
        for l in range(nsteps):
            xi = tr.step()
            if xi > 0:
                num_ref[l] += wt
        np.testing.assert_allclose(q1[t], tr.q, rtol=0, atol=1e-9)
    np.testing.assert_allclose(den, den_ref, rtol=1e-12)
    np.testing.assert_allclose(num, num_ref, rtol=1e-12, atol=1e-12)
    e.close()


def test_children_continue_across_runs():
    """Two runs of 30 steps continue the same children: the kappa numerator
    and the state equal those of one 60-step run, bit for bit."""
    w = workloads.ch4h_children(ntraj=48, nbeads=8, seed=6)
    out = []
    for runs in ([60], [30, 30]):
        e = cam.RPMDEngine(cam.CAM_SURFACE_CH4H, 8, 48, w["dt"], w["masses"], ensemble=cam.CAM_ENS_CHILD,
                           reaction=w["reaction"], seed=2, steps_per_launch=10)
        e.set_state(w["q"], w["p"], w["beta"])
        e.begin_children()
        for n in runs:
            e.run(n)
        num, den = e.kappa(60)
        out.append((num, den, e.get_state()[0]))
        e.close()
    assert np.array_equal(out[0][0], out[1][0]) and out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2])


def test_autotune_leaves_the_state_and_keeps_parity():
    """The autotuner times schedules on a copy: state, kappa sums and step
    counters are untouched, and the schedule it keeps gives the same
    trajectory as the default one to rounding."""
    w = workloads.ch4h_children(ntraj=64, nbeads=16, seed=7)
    res = []
    for tune in (False, True):
        e = cam.RPMDEngine(cam.CAM_SURFACE_CH4H, 16, 64, w["dt"], w["masses"], ensemble=cam.CAM_ENS_CHILD,
                           reaction=w["reaction"], seed=2)
        e.set_state(w["q"], w["p"], w["beta"])
        e.begin_children()
        q0, p0 = e.get_state()
        if tune:
            trials, best = e.autotune(20)
            print("autotune trials", trials, "chosen", best)
            # VALU and MFMA: 2 budgets x 4 launch lengths + split; folded: 2 x 4
            assert len(trials) == 26 and best in trials
            q1, p1 = e.get_state()
            assert np.array_equal(q0, q1) and np.array_equal(p0, p1)
        e.run(60)
        res.append((e.get_state()[0], e.kappa(60)))
        e.close()
    (qa, (na, da)), (qb, (nb_, db)) = res
    assert np.abs(qa - qb).max() <= 1e-11 * np.abs(qa).max()
    assert da == db
    np.testing.assert_allclose(na, nb_, rtol=0, atol=1e-12 * max(1.0, np.abs(na).max()))


@pytest.mark.parametrize("ens", [cam.CAM_ENS_NVE, cam.CAM_ENS_ANDERSEN, cam.CAM_ENS_CHILD])
@pytest.mark.parametrize("surface,nb", [(cam.CAM_SURFACE_CH4H, 32), (cam.CAM_SURFACE_CH4H, 8), (cam.CAM_SURFACE_H3, 16)])
def test_split_schedule_matches_fused(ens, surface, nb):
    """The split schedule (normal-mode kernel + force kernel per step, last
    half kick at the end) follows the fused trajectory to rounding over 60
    steps, including Andersen redraws and the recrossing flags."""
    if surface == cam.CAM_SURFACE_CH4H:
        w = workloads.ch4h_children(ntraj=40, nbeads=nb, seed=8)
        q, p, beta, rx = w["q"], w["p"], w["beta"], w["reaction"]
        rng = np.random.default_rng(8)
        p = rng.normal(size=q.shape) * np.sqrt(w["masses"][None, None, :, None] * nb / beta[:, None, None, None])
    else:
        w = workloads.h3_nve(nbeads=nb)
        q, p, beta, rx = np.stack([w["q"]] * 8), np.stack([w["p"]] * 8), np.full(8, w["beta"]), workloads.h3_reaction()
    out = []
    for sched in (cam.CAM_SCHED_FUSED, cam.CAM_SCHED_SPLIT):
        e = cam.RPMDEngine(surface, nb, q.shape[0], w["dt"], w["masses"], ensemble=ens, andersen_step=7, seed=5,
                           reaction=rx, schedule=sched, steps_per_launch=25)
        e.set_state(q, p, beta)
        if ens == cam.CAM_ENS_CHILD:
            e.begin_children()
        e.run(30)
        e.run(30)
        qq, pp = e.get_state()
        kap = e.kappa(60) if ens == cam.CAM_ENS_CHILD else None
        out.append((qq, pp, kap))
        e.close()
    (q0, p0, k0), (q1, p1, k1) = out
    err = np.abs(q0 - q1).max() / np.abs(q0).max()
    print(f"ens={ens} surface={surface} nb={nb}: split vs fused max rel position difference {err:.2e}")
    assert err <= 1e-11
    assert np.abs(p0 - p1).max() <= 1e-9 * np.abs(p0).max()
    if k0 is not None:
        assert k0[1] == k1[1]
        np.testing.assert_allclose(k0[0], k1[0], rtol=0, atol=1e-12 * max(1.0, np.abs(k0[0]).max()))


def test_mixed_precision_engine_tracks_fp64():
    w = workloads.ch4h_children(ntraj=64, nbeads=16)
    p0 = np.random.default_rng(4).normal(size=w["q"].shape) * 15
    res = []
    for prec in (cam.CAM_PREC_FP64, cam.CAM_PREC_MIXED):
        e = cam.RPMDEngine(cam.CAM_SURFACE_CH4H, 16, 64, w["dt"], w["masses"], precision=prec)
        e.set_state(w["q"], p0, w["beta"])
        e.run(100)
        res.append(e.get_state()[0])
        e.close()
    err = np.abs(res[0] - res[1]).max()
    print(f"mixed vs fp64 after 100 steps: max |dq| = {err:.2e} bohr")
    assert err < 1e-5


@pytest.mark.parametrize("prec", [cam.CAM_PREC_FP64, cam.CAM_PREC_MIXED])
def test_mfma_and_valu_transforms_agree_on_a_ch4h_batch(prec):
    """CH4H children, 32 beads, 40 trajectories (a partial last block): the
    matrix-core normal-mode step and the vector-ALU one stay within rounding
    of each other over 200 steps, and the recrossing flags agree."""
    w = workloads.ch4h_children(ntraj=40, nbeads=32, seed=4)
    res = []
    for nmk in (cam.CAM_NMK_VALU, cam.CAM_NMK_MFMA):
        e = cam.RPMDEngine(cam.CAM_SURFACE_CH4H, 32, 40, w["dt"], w["masses"], precision=prec,
                           ensemble=cam.CAM_ENS_CHILD, reaction=w["reaction"], seed=3, nm_kernel=nmk)
        e.set_state(w["q"], w["p"], w["beta"])
        e.begin_children()
        e.run(200)
        q, p = e.get_state()
        num, den = e.kappa(200)
        res.append((q, p, num, den))
        e.close()
    (q0, p0, n0, d0), (q1, p1, n1, d1) = res
    err = np.abs(q0 - q1).max() / np.abs(q0).max()
    print(f"prec={prec} MFMA vs VALU max rel position difference {err:.2e}")
    assert d0 == d1
    if prec == cam.CAM_PREC_FP64:
        assert err <= 1e-11
        assert np.abs(p0 - p1).max() <= 1e-9 * np.abs(p0).max()
        np.testing.assert_allclose(n0, n1, rtol=0, atol=1e-12 * max(1.0, np.abs(n0).max()))
    else:
        # FP32 forces turn rounding-level position differences into FP32-level ones
        assert err <= 1e-6
