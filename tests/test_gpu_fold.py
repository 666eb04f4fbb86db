"""GPU parity of the folded step kernel (nm_kernel FOLDED): the reference
transform makes every ring polymer symmetric under the bead reflection
j -> nbeads - j, so the kernel carries beads 0..nbeads/2 only.  Checked
against the oracle's verlet (which transforms and evaluates every bead) and
against the unfolded kernels, from states that start asymmetric."""
import numpy as np
import pytest

import caracal_amd as cam
import oracle_lib as ol
from caracal_amd import workloads

pytestmark = pytest.mark.gpu


def _mirror_symmetric(x):
    """x[..., j, :, :] == x[..., nb - j, :, :] for every bead j, bit for bit."""
    nb = x.shape[-3]
    return all(np.array_equal(x[..., j, :, :], x[..., (nb - j) % nb, :, :]) for j in range(nb))


@pytest.mark.parametrize("nb", [4, 6, 8, 12, 16, 24, 32, 48, 64])
def test_folded_h3_nve_trajectory_matches_oracle_100_steps(nb):
    w = workloads.h3_nve(nbeads=nb)
    rng = np.random.default_rng(nb)
    # momenta that are not reflection-symmetric, so the first fold matters
    p0 = w["p"] + rng.normal(size=w["p"].shape) * np.abs(w["p"]).max() * 0.3
    e = cam.RPMDEngine(cam.CAM_SURFACE_H3, nb, 1, w["dt"], w["masses"], nm_kernel=cam.CAM_NMK_FOLDED,
                       steps_per_launch=25)
    e.set_state(w["q"][None], p0[None], w["beta"])
    s = ol.make_sys(1, nb, w["dt"], w["masses"])
    t = ol.OracleTraj(s, w["q"], p0, w["beta"], ensemble=0)
    for block in range(4):
        e.run(25)
        for _ in range(25):
            t.step()
        q, p = e.get_state()
        err = np.abs(q[0] - t.q).max() / np.abs(t.q).max()
        print(f"nb={nb} step={25 * (block + 1)} max rel position error {err:.2e}")
        assert err <= 1e-10
        assert np.abs(p[0] - t.p).max() <= 1e-9 * np.abs(t.p).max()
        assert _mirror_symmetric(q) and _mirror_symmetric(p)
    v, g = e.get_forces()
    assert _mirror_symmetric(g)
    e.close()


@pytest.mark.parametrize("nb", [8, 32])
def test_folded_children_flags_and_kappa_match_oracle(nb):
    ntraj, nsteps = 16, 60
    w = workloads.ch4h_children(ntraj=ntraj, nbeads=nb)
    e = cam.RPMDEngine(cam.CAM_SURFACE_CH4H, nb, ntraj, w["dt"], w["masses"], ensemble=cam.CAM_ENS_CHILD,
                       reaction=w["reaction"], seed=99, steps_per_launch=25, nm_kernel=cam.CAM_NMK_FOLDED)
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


@pytest.mark.parametrize("prec", [cam.CAM_PREC_FP64, cam.CAM_PREC_MIXED])
def test_folded_and_unfolded_agree_on_a_ch4h_batch(prec):
    """CH4H children, 32 beads, 40 trajectories (3 per folded block, so the
    last block holds one): folded and unfolded MFMA steps stay within rounding
    of each other over 200 steps, and the recrossing flags agree."""
    w = workloads.ch4h_children(ntraj=40, nbeads=32, seed=4)
    res = []
    for nmk in (cam.CAM_NMK_MFMA, cam.CAM_NMK_FOLDED):
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
    print(f"prec={prec} folded vs unfolded max rel position difference {err:.2e}")
    assert d0 == d1
    if prec == cam.CAM_PREC_FP64:
        assert err <= 1e-11
        assert np.abs(p0 - p1).max() <= 1e-9 * np.abs(p0).max()
        np.testing.assert_allclose(n0, n1, rtol=0, atol=1e-12 * max(1.0, np.abs(n0).max()))
    else:
        # FP32 forces turn rounding-level position differences into FP32-level ones
        assert err <= 1e-6


@pytest.mark.parametrize("prec", [cam.CAM_PREC_FP64, cam.CAM_PREC_MIXED])
def test_folded_trajectory_slices_give_the_bits_of_small_batches(prec):
    """1600 CH4H children x 32 beads (534 folded blocks: the launches are
    issued as two trajectory slices on two streams): every trajectory ends
    with the bits it has when run in a batch of 40 that is not sliced, on
    both sides of the slice boundary; the recrossing sums follow the unfolded
    kernel to rounding."""
    n, nb, steps = 1600, 32, 120
    w, q0, p0, beta, rx = _batch(cam.CAM_SURFACE_CH4H, nb, n, seed=11)
    kw = dict(precision=prec, ensemble=cam.CAM_ENS_CHILD, reaction=rx, seed=3, steps_per_launch=50)

    def run(sel, nmk):
        e = cam.RPMDEngine(cam.CAM_SURFACE_CH4H, nb, len(sel), w["dt"], w["masses"], nm_kernel=nmk, **kw)
        e.set_state(q0[sel], p0[sel], beta[sel])
        e.begin_children()
        e.run(steps)
        q, p = e.get_state()
        k = e.kappa(steps)
        e.close()
        return q, p, k

    everything = np.arange(n)
    qa, pa, (na, da) = run(everything, cam.CAM_NMK_FOLDED)
    half = (((n + 2) // 3) // 2) * 3               # first trajectory of the second slice
    for lo in (0, half - 20, n - 40):
        sel = np.arange(lo, lo + 40)
        qs, ps, _ = run(sel, cam.CAM_NMK_FOLDED)
        assert np.array_equal(qa[sel], qs) and np.array_equal(pa[sel], ps), lo
    qv, pv, (nv, dv) = run(everything, cam.CAM_NMK_VALU)
    err = np.abs(qa - qv).max() / np.abs(qv).max()
    print(f"prec={prec}: sliced folded vs unfolded rel dq {err:.1e}")
    assert err <= (1e-11 if prec == cam.CAM_PREC_FP64 else 1e-6)
    assert da == dv
    if prec == cam.CAM_PREC_FP64:
        np.testing.assert_allclose(na, nv, rtol=0, atol=1e-12 * max(1.0, np.abs(nv).max()))


@pytest.mark.parametrize("surface,nb,ntraj", [(cam.CAM_SURFACE_H3, 4, 200), (cam.CAM_SURFACE_H3, 64, 20),
                                              (cam.CAM_SURFACE_CH4H, 16, 30), (cam.CAM_SURFACE_CH4H, 64, 3)])
@pytest.mark.parametrize("prec", [cam.CAM_PREC_FP64, cam.CAM_PREC_MIXED])
def test_folded_fused_steps_equal_single_steps(surface, nb, ntraj, prec):
    """One step per launch and 40 steps per launch give the same state bit for
    bit (re-folding a symmetric state is exact), over several blocks."""
    if surface == cam.CAM_SURFACE_H3:
        w = workloads.h3_nve(nbeads=nb)
        q = np.stack([w["q"]] * ntraj)
        beta = np.full(ntraj, w["beta"])
    else:
        w = workloads.ch4h_children(ntraj=ntraj, nbeads=nb, seed=2)
        q, beta = w["q"], w["beta"]
    rng = np.random.default_rng(1)
    p = rng.normal(size=q.shape) * np.sqrt(w["masses"][None, None, :, None] * nb / beta[:, None, None, None])
    out = []
    for spl in (1, 40):
        e = cam.RPMDEngine(surface, nb, ntraj, w["dt"], w["masses"], precision=prec, steps_per_launch=spl,
                           nm_kernel=cam.CAM_NMK_FOLDED)
        e.set_state(q, p, beta)
        e.run(40)
        out.append(e.get_state())
        e.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert _mirror_symmetric(out[0][0]) and _mirror_symmetric(out[0][1])


def _batch(surface, nb, ntraj, seed):
    if surface == cam.CAM_SURFACE_H3:
        w = workloads.h3_nve(nbeads=nb)
        rng = np.random.default_rng(seed)
        q = np.stack([w["q"]] * ntraj) + rng.normal(0, 0.02, (ntraj,) + w["q"].shape)
        beta = np.full(ntraj, w["beta"]) * rng.uniform(0.5, 2.0, ntraj)
        rx = workloads.h3_reaction()
    else:
        w = workloads.ch4h_children(ntraj=ntraj, nbeads=nb, seed=seed)
        q, beta, rx = w["q"], w["beta"], w["reaction"]
        rng = np.random.default_rng(seed)
    p = rng.normal(size=q.shape) * np.sqrt(w["masses"][None, None, :, None] * nb / beta[:, None, None, None])
    return w, q, p, beta, rx


@pytest.mark.parametrize("surface", [cam.CAM_SURFACE_H3, cam.CAM_SURFACE_CH4H])
@pytest.mark.parametrize("nb", [4, 6, 8, 12, 16, 24, 32, 48, 64])
@pytest.mark.parametrize("prec", [cam.CAM_PREC_FP64, cam.CAM_PREC_MIXED])
@pytest.mark.parametrize("ens", [cam.CAM_ENS_NVE, cam.CAM_ENS_CHILD])
def test_every_folded_kernel_is_deterministic_and_tracks_the_unfolded_one(surface, nb, prec, ens):
    """Every folded instantiation, on three blocks of trajectories at
    spread temperatures and with asymmetric starting momenta: two runs agree
    bit for bit, and 5 steps stay within rounding of the unfolded vector-ALU
    kernel (FP32-force level in mixed precision)."""
    L = 1 if surface == cam.CAM_SURFACE_H3 else 4
    tpb = (256 // L) // (nb // 2 + 1)
    ntraj = 2 * tpb + 1
    w, q, p, beta, rx = _batch(surface, nb, ntraj, seed=nb + 7 * ens)
    kw = dict(precision=prec, ensemble=ens, reaction=rx, seed=3, steps_per_launch=5)
    out = []
    for nmk in (cam.CAM_NMK_FOLDED, cam.CAM_NMK_FOLDED, cam.CAM_NMK_VALU):
        e = cam.RPMDEngine(surface, nb, ntraj, w["dt"], w["masses"], nm_kernel=nmk, **kw)
        e.set_state(q, p, beta)
        if ens == cam.CAM_ENS_CHILD:
            e.begin_children()
        e.run(5)
        out.append(e.get_state() + ((e.kappa(5),) if ens == cam.CAM_ENS_CHILD else ()))
        e.close()
    (qa, pa, *ka), (qb, pb, *kb), (qv, pv, *kv) = out
    assert np.array_equal(qa, qb) and np.array_equal(pa, pb)
    err = np.abs(qa - qv).max() / np.abs(qv).max()
    perr = np.abs(pa - pv).max() / np.abs(pv).max()
    print(f"surface={surface} nb={nb} prec={prec} ens={ens}: folded vs unfolded rel dq {err:.1e} dp {perr:.1e}")
    assert err <= (1e-12 if prec == cam.CAM_PREC_FP64 else 1e-9)
    assert perr <= (1e-10 if prec == cam.CAM_PREC_FP64 else 1e-6)
    if ens == cam.CAM_ENS_CHILD:
        assert ka[0][1] == kv[0][1]
        np.testing.assert_array_equal(ka[0][0], kb[0][0])


@pytest.mark.parametrize("nb", [4, 8, 16, 32, 64])
@pytest.mark.parametrize("ens", [cam.CAM_ENS_NVE, cam.CAM_ENS_CHILD])
def test_folded_ch4h_fp64_register_budgets_give_the_same_bits(nb, ens):
    """The FP64 CH5 folded kernel at both register budgets (one and two
    waves per SIMD) over 30 steps in 10-step launches, on three blocks: the
    same arithmetic, so the same bits, and within rounding of the unfolded
    vector-ALU kernel."""
    tpb = 64 // (nb // 2 + 1)
    ntraj = 2 * tpb + 1
    w, q, p, beta, rx = _batch(cam.CAM_SURFACE_CH4H, nb, ntraj, seed=3 * nb + ens)
    kw = dict(precision=cam.CAM_PREC_FP64, ensemble=ens, reaction=rx, seed=5, steps_per_launch=10)
    out = []
    for nmk, wpe in ((cam.CAM_NMK_FOLDED, 1), (cam.CAM_NMK_FOLDED, 2), (cam.CAM_NMK_VALU, 1)):
        e = cam.RPMDEngine(cam.CAM_SURFACE_CH4H, nb, ntraj, w["dt"], w["masses"], nm_kernel=nmk,
                           waves_per_eu=wpe, **kw)
        e.set_state(q, p, beta)
        if ens == cam.CAM_ENS_CHILD:
            e.begin_children()
        e.run(30)
        out.append(e.get_state() + ((e.kappa(30),) if ens == cam.CAM_ENS_CHILD else ()))
        e.close()
    (q1, p1, *k1), (q2, p2, *k2), (qv, pv, *kv) = out
    assert np.array_equal(q1, q2) and np.array_equal(p1, p2)
    err = np.abs(q2 - qv).max() / np.abs(qv).max()
    print(f"nb={nb} ens={ens}: two-wave folded vs unfolded rel dq {err:.1e}")
    assert err <= 1e-12
    if ens == cam.CAM_ENS_CHILD:
        np.testing.assert_array_equal(k1[0][0], k2[0][0])
        assert k1[0][1] == k2[0][1] == kv[0][1]


def test_folded_rejects_what_it_does_not_cover():
    w = workloads.h3_nve(nbeads=8)
    for kw in (dict(ensemble=cam.CAM_ENS_ANDERSEN, andersen_step=5), dict(remove_com=True),
               dict(nm_mode=cam.CAM_NM_ORTHOGONAL), dict(schedule=cam.CAM_SCHED_SPLIT)):
        with pytest.raises(cam.CamError, match="FOLDED"):
            cam.RPMDEngine(cam.CAM_SURFACE_H3, 8, 1, w["dt"], w["masses"], nm_kernel=cam.CAM_NMK_FOLDED, **kw)
    w = workloads.h3_nve(nbeads=2)
    with pytest.raises(cam.CamError, match="FOLDED"):
        cam.RPMDEngine(cam.CAM_SURFACE_H3, 2, 1, w["dt"], w["masses"], nm_kernel=cam.CAM_NMK_FOLDED)

