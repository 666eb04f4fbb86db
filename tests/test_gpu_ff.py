"""GPU parity of the QMDFF-form box: energy kernel, its hand-written adjoint
(forces, parameter gradients, force-matching contraction) against the oracle
and against central differences of the GPU energy kernel (h = 1e-5), the
torch.autograd wrapper, and the box ring-polymer engine."""
import numpy as np
import pytest

import caracal_amd as cam
import oracle_lib as ol
from caracal_amd import ff as cff
from caracal_amd import ffbox

pytestmark = pytest.mark.gpu


def _box(nmol, box_a, seed, rc_a, jitter=0.05):
    s = ffbox.box_system(nmol=nmol, box_angstrom=box_a, seed=seed)
    s["x"] = s["x"] + np.random.default_rng(seed).normal(0, jitter, s["x"].shape)
    return s, cff.ForceField(s, rc_angstrom=rc_a), ol.FFOracle(s, rc_angstrom=rc_a)


def _beads(s, nb, seed, sigma=0.05):
    rng = np.random.default_rng(seed)
    return s["x"][None] + rng.normal(0, sigma, (nb,) + s["x"].shape)


def test_energy_and_forces_match_oracle_full_size_box():
    import torch
    s, ff, fo = _box(1000, 31.0, 2, 10.0)
    th = ffbox.theta_ref()
    xb = _beads(s, 2, 7)
    x = torch.tensor(xb, device="cuda")
    e, g, _, _ = ff.forces(th, x)
    e = e.cpu().numpy()
    g = g.cpu().numpy()
    for b in range(2):
        er, gr, _, _ = fo.eval(th, xb[b])
        assert abs(e[b] - er) < 1e-10 * abs(er)
        np.testing.assert_allclose(g[b], gr, rtol=0, atol=1e-10 * np.abs(gr).max())
    em = ff.energy(th, x, precision=cam.CAM_PREC_MIXED).cpu().numpy()
    _, gm, _, _ = ff.forces(th, x, precision=cam.CAM_PREC_MIXED)
    gm = gm.cpu().numpy()
    # mixed: FP32 pair arithmetic, compensated accumulation
    fe = np.max(np.abs(gm - g).reshape(-1, 3).max(1)) / np.abs(g).max()
    print(f"box mixed-precision max force error relative to max |F|: {fe:.2e}")
    assert fe < 1e-5
    assert np.max(np.abs(em - e) / np.abs(e)) < 1e-5


def test_adjoint_matches_central_differences_of_energy_kernel():
    import torch
    s, ff, _ = _box(64, 12.0, 3, 5.9)
    th = ffbox.theta_ref()
    x = torch.tensor(_beads(s, 1, 1), device="cuda")
    _, g, dth, _ = ff.forces(th, x, want_params=True)
    g, dth = g.cpu().numpy()[0], dth.cpu().numpy()[0]
    h = 1e-5
    rng = np.random.default_rng(2)
    for idx in rng.choice(x.numel(), 30, replace=False):
        xp, xm = x.clone().view(-1), x.clone().view(-1)
        xp[idx] += h
        xm[idx] -= h
        fd = (ff.energy(th, xp.view_as(x)).item() - ff.energy(th, xm.view_as(x)).item()) / (2 * h)
        assert abs(fd - g.ravel()[idx]) < 2e-8 + 2e-6 * abs(fd)
    for p in range(10):
        hp = h * max(1.0, abs(th[p]))
        tp, tm = th.copy(), th.copy()
        tp[p] += hp
        tm[p] -= hp
        fd = (ff.energy(tp, x).item() - ff.energy(tm, x).item()) / (2 * hp)
        assert abs(fd - dth[p]) < 1e-6 * max(1.0, abs(fd)), (p, fd, dth[p])


def test_parameter_gradients_match_oracle():
    import torch
    s, ff, fo = _box(64, 12.0, 5, 5.9)
    th = ffbox.theta_ref() * (1 + 0.05 * np.random.default_rng(0).normal(size=10))
    xb = _beads(s, 3, 4)
    v = np.random.default_rng(9).normal(size=xb.shape)
    x, vt = torch.tensor(xb, device="cuda"), torch.tensor(v, device="cuda")
    e, g, dth, dvth = ff.forces(th, x, v=vt, want_params=True)
    for b in range(3):
        er, gr, dr, dvr = fo.eval(th, xb[b], v=v[b])
        np.testing.assert_allclose(dth.cpu().numpy()[b], dr, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(dvth.cpu().numpy()[b], dvr, rtol=1e-10, atol=1e-12)


def test_autograd_force_matching_gradient():
    import torch
    s, ff, _ = _box(64, 12.0, 6, 5.9)
    fn = cff.make_autograd(ff)
    x = torch.tensor(_beads(s, 2, 3), device="cuda")
    th_ref = torch.tensor(ffbox.theta_ref(), device="cuda")
    with torch.no_grad():
        _, f_ref = fn(x, th_ref)
    th = (th_ref * 1.05).clone().requires_grad_(True)
    e, f = fn(x, th)
    loss = ((f - f_ref) ** 2).sum() + 0.1 * e.sum()
    loss.backward()
    grad = th.grad.cpu().numpy()
    for p in range(10):
        h = 1e-6 * max(1.0, abs(th[p].item()))
        tp, tm = th.detach().clone(), th.detach().clone()
        tp[p] += h
        tm[p] -= h
        with torch.no_grad():
            ep, fp = fn(x, tp)
            em, fm = fn(x, tm)
            lp = ((fp - f_ref) ** 2).sum() + 0.1 * ep.sum()
            lm = ((fm - f_ref) ** 2).sum() + 0.1 * em.sum()
        fd = ((lp - lm) / (2 * h)).item()
        assert abs(fd - grad[p]) < 1e-5 * max(1.0, abs(fd)), (p, fd, grad[p])


def _box_start(nmol, box_a, seed, nb, kelvin=300.0):
    s = ffbox.box_system(nmol=nmol, box_angstrom=box_a, seed=seed)
    m = ffbox.masses(s["type"])
    beta = float(cam.beta_of(kelvin))
    rng = np.random.default_rng(1)
    q = s["x"][None] + rng.normal(0, 0.02, (nb,) + s["x"].shape)
    p = rng.normal(size=q.shape) * np.sqrt(m[None, :, None] * nb / beta)
    return s, m, beta, q, p


@pytest.mark.parametrize("nm_mode", [cam.CAM_NM_REFERENCE, cam.CAM_NM_ORTHOGONAL])
def test_box_engine_matches_oracle_trajectory(nm_mode):
    """125 molecules, 8 beads, 7.5 angstrom cutoff, FP64: 40 steps of the GPU
    box engine against the oracle's ring polymer (oracle force field + the
    free ring-polymer step); positions and the energy terms agree."""
    import box_oracle as bo
    nb = 8
    s, m, beta, q, p = _box_start(125, 15.5, 8, nb)
    ff = cff.ForceField(s, rc_angstrom=7.5)
    dt = 0.1 * cam.FS_TO_AU
    eng = cff.BoxEngine(ff, nb, dt, m, beta, ffbox.theta_ref(), precision=cam.CAM_PREC_FP64, nm_mode=nm_mode)
    eng.set_state(q, p)
    t = bo.BoxRPMD(ol.FFOracle(s, rc_angstrom=7.5), ffbox.theta_ref(), q, p, dt, m, beta,
                   orthogonal=nm_mode == cam.CAM_NM_ORTHOGONAL)
    for block in range(4):
        eng.run(10)
        for _ in range(10):
            t.step()
        qg, pg = eng.get_state()
        err = np.abs(qg - t.q).max() / np.abs(t.q).max()
        print(f"nm_mode={nm_mode} step={10 * (block + 1)} max rel position error {err:.2e}")
        assert err < 1e-10
        assert np.abs(pg - t.p).max() < 1e-8 * np.abs(t.p).max()
    ep, ek, es = eng.energies()
    vo, eko, eso = t.energies()
    np.testing.assert_allclose(ep, vo, rtol=1e-9)
    np.testing.assert_allclose([ek, es], [eko, eso], rtol=1e-9)
    eng.close()


def test_box_engine_nve_conserves_energy():
    """Orthogonal normal modes on a 27-molecule cluster (12 angstrom across)
    in a 31 angstrom box with a 15 angstrom cutoff, so that no pair crosses the
    cutoff (where the reference's energy jumps by q_i q_j erfc(a rc)/rc): the
    ring-polymer energy error stays bounded and is second order in dt."""
    from caracal_amd.units import BOHR_ANGSTROM
    nb = 8
    s = ffbox.box_system(nmol=27, box_angstrom=9.3, seed=8)
    s["box"] = 31.0 / BOHR_ANGSTROM
    ff = cff.ForceField(s, rc_angstrom=15.0)
    m = ffbox.masses(s["type"])
    beta = float(cam.beta_of(300.0))
    dev = {}
    for dtfs, nsteps in ((0.1, 200), (0.05, 400)):
        rng = np.random.default_rng(1)
        q = s["x"][None] + rng.normal(0, 0.02, (nb,) + s["x"].shape)
        p = rng.normal(size=q.shape) * np.sqrt(m[None, :, None] * nb / beta)
        eng = cff.BoxEngine(ff, nb, dtfs * cam.FS_TO_AU, m, beta, ffbox.theta_ref(), precision=cam.CAM_PREC_FP64,
                            nm_mode=cam.CAM_NM_ORTHOGONAL)
        eng.set_state(q, p)

        def h():
            ep, ek, es = eng.energies()
            return ep.sum() + ek + es
        h0 = h()
        hs = []
        for _ in range(20):
            eng.run(nsteps // 20)
            hs.append(h() - h0)
        eng.close()
        dev[dtfs] = np.abs(np.array(hs))
        print(f"dt={dtfs} fs max |dH| {dev[dtfs].max():.3e}")
    assert dev[0.1].max() < 5e-3
    ratio = dev[0.1].mean() / dev[0.05].mean()
    assert 3.0 < ratio < 5.0


def test_force_matching_recovers_parameters():
    """Force matching from the reference parameters perturbed by 10% (seed
    3), through the adjoint: the loss gradient at the start agrees with
    central differences (h = 1e-5 relative) of the loss itself, and up to 50
    L-BFGS iterations bring the loss down by five orders of magnitude and the
    seven force-determined parameters back to within 1e-3 (the three
    dispersion coefficients, which the forces do not fix, stay where they
    started)."""
    import torch
    from caracal_amd import fit
    s, ff, _ = _box(125, 15.5, 9, 7.5)
    x = torch.tensor(fit.box_beads(s, 8, seed=4), device="cuda")
    th_ref = ffbox.theta_ref()
    th0 = fit.perturbed_theta(th_ref, 0.10, seed=3)
    fn = cff.make_autograd(ff)
    scale = torch.tensor(th_ref, device="cuda")
    with torch.no_grad():
        _, f_ref = fn(x, scale)
    norm = float((f_ref * f_ref).sum())

    def loss_of(u):
        _, f = fn(x, scale * (1.0 + u))
        return ((f - f_ref) ** 2).sum() / norm

    u = torch.tensor(th0 / th_ref - 1.0, device="cuda", requires_grad=True)
    loss_of(u).backward()
    grad = u.grad.cpu().numpy()
    for p in range(10):
        h = 1e-5
        up, um = u.detach().clone(), u.detach().clone()
        up[p] += h
        um[p] -= h
        with torch.no_grad():
            fd = float((loss_of(up) - loss_of(um)) / (2 * h))
        assert abs(fd - grad[p]) <= 1e-6 + 1e-4 * abs(fd), (p, fd, grad[p])
    res = fit.force_match(ff, x, th_ref, th0, steps=50)
    print(f"force matching: loss {res['loss'][0]:.3e} -> {res['loss'][-1]:.3e} in {res['iterations']} L-BFGS "
          f"iterations; force-determined parameters {res['rel_err_forces0']:.3f} -> {res['rel_err_forces']:.2e}, "
          f"dispersion {res['rel_err_dispersion0']:.3f} -> {res['rel_err_dispersion']:.3f}")
    assert res["loss"][-1] < 1e-5 * res["loss"][0]
    # the seven parameters the forces determine come back; the dispersion
    # coefficients (Gauss-Newton curvature ~1e-9 of the largest) stay put
    assert res["rel_err_forces"] < 1e-3
    assert res["rel_err_dispersion"] < res["rel_err_dispersion0"] + 1e-3


@pytest.mark.parametrize("prec", [cam.CAM_PREC_FP64, cam.CAM_PREC_MIXED])
@pytest.mark.parametrize("sigma", [0.05, 0.8])
def test_neighbour_list_equals_all_pairs_bitwise(prec, sigma):
    """The neighbour-list pair loop (default) and the all-pairs loop give the
    same energies, forces, parameter gradients and force-matching contraction
    bit for bit, for ring polymers of 8 beads spread by sigma bohr around the
    box (sigma 0.8: the bead spread enters the list radius), and over 30 steps
    of the box engine."""
    import torch
    s, ff, _ = _box(1000, 31.0, 4, 10.0)
    th = ffbox.theta_ref()
    x = torch.tensor(_beads(s, 8, 5, sigma=sigma), device="cuda")
    v = torch.tensor(np.random.default_rng(6).normal(size=tuple(x.shape)), device="cuda")
    out = {}
    for mode in (cam._abi.CAM_FF_PAIRS_ALL, cam._abi.CAM_FF_PAIRS_LIST):
        ff.set_pair_mode(mode)
        e, g, dth, dvth = ff.forces(th, x, v=v, want_params=True, precision=prec)
        out[mode] = [t.cpu().numpy() for t in (e, g, dth, dvth)] + [ff.energy(th, x, precision=prec).cpu().numpy()]
    for a, b in zip(out[cam._abi.CAM_FF_PAIRS_ALL], out[cam._abi.CAM_FF_PAIRS_LIST]):
        assert np.array_equal(a, b)
    # the box engine's steps (forces every step, list rebuilt every step)
    m = ffbox.masses(s["type"])
    beta = float(cam.beta_of(300.0))
    rng = np.random.default_rng(7)
    q = _beads(s, 8, 8, sigma=sigma)
    p = rng.normal(size=q.shape) * np.sqrt(m[None, :, None] * 8 / beta)
    res = []
    for mode in (cam._abi.CAM_FF_PAIRS_ALL, cam._abi.CAM_FF_PAIRS_LIST):
        ff.set_pair_mode(mode)
        eng = cff.BoxEngine(ff, 8, 0.1 * cam.FS_TO_AU, m, beta, th, precision=prec, andersen_step=10, seed=2)
        eng.set_state(q, p)
        eng.run(30)
        res.append(eng.get_state())
        eng.close()
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
    ff.close()


@pytest.mark.parametrize("prec", [cam.CAM_PREC_FP64, cam.CAM_PREC_MIXED])
def test_cell_sorted_rows_equal_all_pairs_bitwise_large_box(prec):
    """A box wide enough for the cell search (4000 molecules, 12000 atoms,
    49.2 angstrom, four cells per side): rows and all pairs give the same
    bits, and the FP64 energy matches the oracle on one bead."""
    import torch
    s, ff, fo = _box(4000, 49.2, 8, 10.0)
    th = ffbox.theta_ref()
    xb = _beads(s, 4, 3, sigma=0.1)
    x = torch.tensor(xb, device="cuda")
    v = torch.tensor(np.random.default_rng(4).normal(size=tuple(x.shape)), device="cuda")
    out = {}
    for mode in (cam._abi.CAM_FF_PAIRS_ALL, cam._abi.CAM_FF_PAIRS_LIST):
        ff.set_pair_mode(mode)
        e, g, dth, dvth = ff.forces(th, x, v=v, want_params=True, precision=prec)
        out[mode] = [t.cpu().numpy() for t in (e, g, dth, dvth)]
    for a, b in zip(out[cam._abi.CAM_FF_PAIRS_ALL], out[cam._abi.CAM_FF_PAIRS_LIST]):
        assert np.array_equal(a, b)
    if prec == cam.CAM_PREC_FP64:
        er, gr, _, _ = fo.eval(th, xb[1])
        assert abs(out[0][0][1] - er) < 1e-10 * abs(er)
        np.testing.assert_allclose(out[0][1][1], gr, rtol=0, atol=1e-10 * np.abs(gr).max())
    ff.close()


def test_rows_rebuilt_only_past_half_skin_and_stay_exact():
    """A sequence of calls on one force field: small moves (every bead by
    0.35 bohr, under half the 0.8 bohr skin: the rows of the first call are
    kept), then some atoms by 0.6 bohr (past it: rebuilt), then a new bead
    count.  Every call equals the all-pairs loop bit for bit and the oracle
    to 1e-10."""
    import torch
    s, ff, fo = _box(1000, 31.0, 4, 10.0)
    th = ffbox.theta_ref()
    rng = np.random.default_rng(12)
    x0 = _beads(s, 4, 5, sigma=0.1)

    def moved(x, frac, length):
        d = rng.normal(size=x.shape)
        d *= length / np.linalg.norm(d, axis=-1, keepdims=True)
        pick = rng.random(x.shape[:2]) < frac
        return x + d * pick[..., None]

    seq = [x0, moved(x0, 1.0, 0.35), moved(x0, 0.02, 0.6), _beads(s, 2, 9, sigma=0.1)]
    for k, xb in enumerate(seq):
        x = torch.tensor(xb, device="cuda")
        res = []
        for mode in (cam._abi.CAM_FF_PAIRS_LIST, cam._abi.CAM_FF_PAIRS_ALL):
            ff.set_pair_mode(mode)
            e, g, _, _ = ff.forces(th, x)
            res.append((e.cpu().numpy(), g.cpu().numpy()))
        assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1]), k
        er, gr, _, _ = fo.eval(th, xb[-1])
        assert abs(res[0][0][-1] - er) < 1e-10 * abs(er), k
        np.testing.assert_allclose(res[0][1][-1], gr, rtol=0, atol=1e-10 * np.abs(gr).max())
    ff.close()


def test_box_engine_nhc_matches_oracle():
    """The box with the Nose-Hoover chain (thermostat 2): 125 molecules,
    8 beads, momenta 1.5 times too hot; 30 steps against the oracle's ring
    polymer with the same chain; set_state restarts the chain."""
    import box_oracle as bo
    nb = 8
    s, m, beta, q, p = _box_start(125, 15.5, 8, nb)
    p = 1.5 * p
    ff = cff.ForceField(s, rc_angstrom=7.5)
    dt = 0.1 * cam.FS_TO_AU
    nose_q = 20.0 * dt
    eng = cff.BoxEngine(ff, nb, dt, m, beta, ffbox.theta_ref(), precision=cam.CAM_PREC_FP64)
    eng.set_nhc(nose_q)
    eng.set_state(q, p)
    t = bo.BoxRPMD(ol.FFOracle(s, rc_angstrom=7.5), ffbox.theta_ref(), q, p, dt, m, beta, nose_q=nose_q)
    eng.run(30)
    for _ in range(30):
        t.step()
    qg, pg = eng.get_state()
    err = np.abs(qg - t.q).max() / np.abs(t.q).max()
    perr = np.abs(pg - t.p).max() / np.abs(t.p).max()
    print(f"box NHC: rel dq {err:.1e} dp {perr:.1e}; chain v1 {t.vnh[0]:.3e}")
    assert err < 1e-10 and perr < 1e-8
    assert abs(t.vnh[0]) > 0
    eng.set_state(q, p)
    eng.run(30)
    q2, _ = eng.get_state()
    assert np.array_equal(q2, qg)
    with pytest.raises(cam.CamError, match="andersen"):
        e2 = cff.BoxEngine(ff, nb, dt, m, beta, ffbox.theta_ref(), andersen_step=5)
        e2.set_nhc(nose_q)
    eng.close()


def test_adjoint_and_force_matching_at_3000_atoms_64_beads():
    """The force-matching configuration at its size: 1000 molecules (3000
    atoms) in the 31 angstrom box, 64 beads, 10 angstrom cutoff.  dE/dtheta
    and d(v . dE/dx)/dtheta of two beads (0 and 63) match the oracle to
    1e-10, and L-BFGS from the 10% perturbation recovers the seven
    force-determined parameters to 1e-3."""
    import torch
    from caracal_amd import fit
    s, xb = fit.default_problem(nbeads=64)
    ff = cff.ForceField(s, rc_angstrom=10.0)
    fo = ol.FFOracle(s, rc_angstrom=10.0)
    th = ffbox.theta_ref() * (1 + 0.05 * np.random.default_rng(1).normal(size=10))
    v = np.random.default_rng(2).normal(size=xb.shape)
    x, vt = torch.tensor(xb, device="cuda"), torch.tensor(v, device="cuda")
    e, g, dth, dvth = ff.forces(th, x, v=vt, want_params=True)
    for b in (0, 63):
        er, gr, dr, dvr = fo.eval(th, xb[b], v=v[b])
        assert abs(e.cpu().numpy()[b] - er) < 1e-10 * abs(er)
        np.testing.assert_allclose(dth.cpu().numpy()[b], dr, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(dvth.cpu().numpy()[b], dvr, rtol=1e-10, atol=1e-12)
    th_ref = ffbox.theta_ref()
    res = fit.force_match(ff, x, th_ref, fit.perturbed_theta(th_ref, 0.10, seed=3), steps=50)
    print(f"3000 x 64 force matching: loss {res['loss'][0]:.3e} -> {res['loss'][-1]:.3e} in {res['iterations']} "
          f"iterations ({res['s_per_eval'] * 1e3:.1f} ms per evaluation); force-determined parameters "
          f"{res['rel_err_forces0']:.3f} -> {res['rel_err_forces']:.2e}")
    assert res["loss"][-1] < 1e-5 * res["loss"][0]
    assert res["rel_err_forces"] < 1e-3
    assert res["rel_err_dispersion"] < res["rel_err_dispersion0"] + 1e-3
    ff.close()
