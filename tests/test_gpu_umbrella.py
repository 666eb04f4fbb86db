"""GPU parity of umbrella-sampling windows (ensemble UMBRELLA): biased
equilibration then unbiased-start sampling, one window per trajectory, with
Andersen draws, against the oracle; the xi statistics; the window check."""
import numpy as np
import pytest

import caracal_amd as cam
import oracle_lib as ol
from caracal_amd import workloads

pytestmark = pytest.mark.gpu


def _windows(surface, nb, nwin, seed):
    if surface == cam.CAM_SURFACE_H3:
        w = workloads.h3_nve(nbeads=nb)
        rng = np.random.default_rng(seed)
        q = np.stack([w["q"]] * nwin) + rng.normal(0, 0.01, (nwin,) + w["q"].shape)
        beta = np.full(nwin, w["beta"])
        rx = workloads.h3_reaction()
    else:
        w = workloads.ch4h_children(ntraj=nwin, nbeads=nb, seed=seed)
        q, beta, rx = w["q"], w["beta"], w["reaction"]
        rng = np.random.default_rng(seed)
    p = rng.normal(size=q.shape) * np.sqrt(w["masses"][None, None, :, None] * nb / beta[:, None, None, None])
    xi_w = np.linspace(0.9, 1.06, nwin)
    k_w = np.linspace(0.02, 0.2, nwin)
    return w, q, p, beta, rx, xi_w, k_w


@pytest.mark.parametrize("surface,nb,nmk,schedule,remove_com", [
    (cam.CAM_SURFACE_CH4H, 8, cam.CAM_NMK_VALU, cam.CAM_SCHED_AUTO, False),   # fused: bias inside the step kernel
    (cam.CAM_SURFACE_CH4H, 8, cam.CAM_NMK_VALU, cam.CAM_SCHED_AUTO, True),    # ... with transrot every step
    (cam.CAM_SURFACE_CH4H, 32, cam.CAM_NMK_MFMA, cam.CAM_SCHED_AUTO, False),  # MFMA normal modes: split schedule
    (cam.CAM_SURFACE_CH4H, 32, cam.CAM_NMK_MFMA, cam.CAM_SCHED_AUTO, True),
    (cam.CAM_SURFACE_H3, 16, cam.CAM_NMK_VALU, cam.CAM_SCHED_AUTO, False),
    (cam.CAM_SURFACE_H3, 16, cam.CAM_NMK_VALU, cam.CAM_SCHED_AUTO, True),
    (cam.CAM_SURFACE_H3, 16, cam.CAM_NMK_VALU, cam.CAM_SCHED_SPLIT, False),
    (cam.CAM_SURFACE_H3, 16, cam.CAM_NMK_VALU, cam.CAM_SCHED_SPLIT, True),
    (cam.CAM_SURFACE_H3, 1, cam.CAM_NMK_VALU, cam.CAM_SCHED_AUTO, False),     # one bead: its own centroid
    (cam.CAM_SURFACE_CH4H, 4, cam.CAM_NMK_VALU, cam.CAM_SCHED_AUTO, False),
    (cam.CAM_SURFACE_H3, 6, cam.CAM_NMK_VALU, cam.CAM_SCHED_AUTO, False),      # bead counts that leave
    (cam.CAM_SURFACE_CH4H, 12, cam.CAM_NMK_VALU, cam.CAM_SCHED_AUTO, True)])   # slots of a block unused
def test_umbrella_windows_match_oracle(surface, nb, nmk, schedule, remove_com):
    nwin, ast, seed, n_equi, n_samp = 6, 5, 21, 20, 30
    w, q, p, beta, rx, xi_w, k_w = _windows(surface, nb, nwin, seed=nb)
    e = cam.RPMDEngine(surface, nb, nwin, w["dt"], w["masses"], ensemble=cam.CAM_ENS_UMBRELLA, reaction=rx,
                       andersen_step=ast, seed=seed, nm_kernel=nmk, schedule=schedule, steps_per_launch=7,
                       remove_com=remove_com)
    e.set_state(q, p, beta)
    e.set_windows(xi_w, k_w)
    e.begin_umbrella(bias=True)
    s = ol.make_sys(surface, nb, w["dt"], w["masses"])
    refs = []
    for t in range(nwin):
        orx = ol.make_rx(rx)
        orx.xi_ideal = xi_w[t]
        tr = ol.OracleTraj(s, q[t], p[t], float(beta[t]), rx=orx, ensemble=4, andersen_step=ast, seed=seed, traj=t,
                           remove_com=remove_com)
        tr.umbrella_begin(k_w[t], bias=True)
        refs.append(tr)
    _, g0 = e.get_forces()
    for t in range(nwin):
        np.testing.assert_allclose(g0[t], refs[t].g, rtol=0, atol=1e-12 * np.abs(refs[t].g).max())
    e.run(n_equi)
    for tr in refs:
        for _ in range(n_equi):
            tr.step()
    e.begin_umbrella(bias=False)
    for t, tr in enumerate(refs):
        tr.umbrella_begin(k_w[t], bias=False)
    e.run(n_samp)
    for tr in refs:
        for _ in range(n_samp):
            tr.step()
    qe, pe = e.get_state()
    avg, var = e.umbrella_stats()
    for t, tr in enumerate(refs):
        err = np.abs(qe[t] - tr.q).max() / np.abs(tr.q).max()
        perr = np.abs(pe[t] - tr.p).max() / np.abs(tr.p).max()
        a_ref = tr.xi_sum / n_samp
        v_ref = tr.xi_sq / n_samp - a_ref * a_ref
        print(f"surface={surface} nb={nb} transrot={remove_com} window {xi_w[t]:.3f}: rel dq {err:.1e} dp {perr:.1e}; "
              f"<xi> {avg[t]:.6f} (oracle {a_ref:.6f}) var {var[t]:.3e}")
        assert err <= 1e-10 and perr <= 1e-9
        assert abs(avg[t] - a_ref) <= 1e-12
        assert abs(var[t] - v_ref) <= 1e-12
    assert not e.observables()["flags"].any()
    e.close()


def test_umbrella_window_check_and_errors():
    w, q, p, beta, rx, xi_w, k_w = _windows(cam.CAM_SURFACE_H3, 8, 4, seed=5)
    e = cam.RPMDEngine(cam.CAM_SURFACE_H3, 8, 4, w["dt"], w["masses"], ensemble=cam.CAM_ENS_UMBRELLA, reaction=rx)
    e.set_state(q, p, beta)
    with pytest.raises(cam.CamError, match="begin_umbrella"):
        e.run(1)
    e.set_windows(np.full(4, 1.0), np.full(4, 0.1))
    e.begin_umbrella(bias=True)
    with pytest.raises(cam.CamError, match="no window steps"):
        e.umbrella_stats()
    # windows at xi = 1 from structures near xi = 1: a tolerance of 0.5 passes, 1e-6 fails from step 3 on
    e.begin_umbrella(bias=True, xi_tol=0.5, check_after=0)
    e.run(5)
    assert not e.observables()["flags"].any()
    e.begin_umbrella(bias=True, xi_tol=1e-6, check_after=2)
    e.run(2)
    assert not e.observables()["flags"].any()
    e.run(1)
    assert (e.observables()["flags"] == 4).all()
    e.close()
    with pytest.raises(cam.CamError, match="remove_com"):
        cam.RPMDEngine(cam.CAM_SURFACE_H3, 8, 1, w["dt"], w["masses"], ensemble=cam.CAM_ENS_PARENT, reaction=rx,
                       remove_com=True)


@pytest.mark.parametrize("surface,nb,schedule", [
    (cam.CAM_SURFACE_CH4H, 8, cam.CAM_SCHED_AUTO),     # fused window step
    (cam.CAM_SURFACE_H3, 16, cam.CAM_SCHED_SPLIT),     # split: umbrella_end_kernel
])
def test_umbrella_histograms_count_the_oracle_samples(surface, nb, schedule):
    """Device histograms of the sampling steps (for WHAM): every window's
    counts equal the oracle's xi of the same steps binned as
    nint((xi - xi_min) / bin_size) on 1..nbins, leaving out samples within
    1e-6 of a bin edge (rounding of near-equal xi may put them either side);
    samples outside the range are not counted; begin_umbrella clears them."""
    nwin, ast, seed, n_equi, n_samp = 6, 5, 33, 15, 40
    w, q, p, beta, rx, xi_w, k_w = _windows(surface, nb, nwin, seed=nb + 1)
    xi_min, xi_max, nbins = 0.86, 1.10, 241
    bs = (xi_max - xi_min) / (nbins - 1)
    e = cam.RPMDEngine(surface, nb, nwin, w["dt"], w["masses"], ensemble=cam.CAM_ENS_UMBRELLA, reaction=rx,
                       andersen_step=ast, seed=seed, schedule=schedule, steps_per_launch=9)
    e.set_state(q, p, beta)
    e.set_windows(xi_w, k_w)
    e.set_umbrella_histogram(nbins, xi_min, bs)
    e.begin_umbrella(bias=True)
    e.run(n_equi)
    assert e.umbrella_histogram().sum() > 0
    e.begin_umbrella(bias=False)
    assert e.umbrella_histogram().sum() == 0
    e.run(n_samp)
    hist = e.umbrella_histogram()
    avg, _ = e.umbrella_stats()
    s = ol.make_sys(surface, nb, w["dt"], w["masses"])
    hinv = 1.0 / bs
    for t in range(nwin):
        orx = ol.make_rx(rx)
        orx.xi_ideal = xi_w[t]
        tr = ol.OracleTraj(s, q[t], p[t], float(beta[t]), rx=orx, ensemble=4, andersen_step=ast, seed=seed, traj=t)
        tr.umbrella_begin(k_w[t], bias=True)
        for _ in range(n_equi):
            tr.step()
        tr.umbrella_begin(k_w[t], bias=False)
        xs = np.array([tr.step() for _ in range(n_samp)])
        assert abs(avg[t] - xs.mean()) <= 1e-12
        u = (xs - xi_min) * hinv
        j = np.rint(u)
        clear = np.abs(np.abs(u - np.floor(u)) - 0.5) > 1e-6
        inside = (j >= 1) & (j <= nbins)
        ref = np.zeros(nbins)
        np.add.at(ref, (j[inside & clear] - 1).astype(int), 1.0)
        near = np.zeros(nbins, dtype=bool)
        for jj in np.floor(u[~clear]):
            near[[int(k) for k in (jj - 1, jj) if 0 <= k < nbins]] = True
        assert np.array_equal(hist[t][~near], ref[~near]), t
        assert hist[t].sum() == (inside).sum() or (~clear).any()
        print(f"surface={surface} window {xi_w[t]:.3f}: {int(hist[t].sum())} of {n_samp} samples counted, "
              f"{int(inside.sum())} inside the range")
    e.set_umbrella_histogram(0, 0.0, 1.0)
    e.close()
