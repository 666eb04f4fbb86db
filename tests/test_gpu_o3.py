"""GPU parity of the O + O2 surface (src/egrad_o3.f, Varga, Paukku and Truhlar) behind the
drop-in egrad contract and inside the trajectory engine, against the oracle:
forces over clouds and the special geometries (O2 with a far O, linear and
compact shapes), NVE trajectories on the normal-mode kernels and bead
counts, recrossing children of the O + O2 exchange with kappa, and an
umbrella window.

The device takes the partial by each MEG variable from a pass of the
polynomial recursion with that variable first (the polynomials are
symmetric), the oracle from one pass of dual numbers; energies near the
saddle cancel terms of up to 1e5 kcal/mol, so they are compared to 1e-12 of
max(1, |V|) and the forces to 1e-11 of their largest magnitude."""
import numpy as np
import pytest
import torch

import caracal_amd as cam
import oracle_lib as ol
from caracal_amd import workloads
from test_oracle_o3 import _cloud, _special

pytestmark = pytest.mark.gpu


def test_o3_forces_match_oracle():
    q = np.concatenate([_cloud(3000, 0.1, 0), _cloud(1000, 0.4, 1), _special()])
    v, g = cam.egrad_o3(q)
    vr, gr = ol.egrad_o3(q)
    err_v = np.abs(v - vr).max() / max(1.0, np.abs(vr).max())
    err_g = np.abs(g - gr).max() / max(1.0, np.abs(gr).max())
    print(f"O3 {len(q)} beads: max rel |dV| {err_v:.1e}, max rel |dg| {err_g:.1e}")
    assert err_v <= 1e-12
    assert err_g <= 1e-11
    qd = torch.tensor(q, device="cuda")
    vd = torch.empty(len(q), dtype=torch.float64, device="cuda")
    gd = torch.empty_like(qd)
    cam.egrad_dev(cam.CAM_SURFACE_O3, qd.data_ptr(), len(q), vd.data_ptr(), gd.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(vd.cpu().numpy(), v) and np.array_equal(gd.cpu().numpy(), g)
    with pytest.raises(cam.CamError, match="FP64 only"):
        cam.egrad_dev(cam.CAM_SURFACE_O3, qd.data_ptr(), len(q), vd.data_ptr(), gd.data_ptr(),
                      precision=cam.CAM_PREC_MIXED)


@pytest.mark.parametrize("nb,nmk", [(1, cam.CAM_NMK_AUTO), (6, cam.CAM_NMK_AUTO), (16, cam.CAM_NMK_VALU),
                                     (32, cam.CAM_NMK_MFMA), (32, cam.CAM_NMK_FOLDED)])
def test_o3_nve_trajectory_matches_oracle(nb, nmk):
    w = workloads.o3_nve(nbeads=nb)
    e = cam.RPMDEngine(cam.CAM_SURFACE_O3, nb, 1, w["dt"], w["masses"], nm_kernel=nmk, steps_per_launch=20)
    e.set_state(w["q"][None], w["p"][None], w["beta"])
    s = ol.make_sys(5, nb, w["dt"], w["masses"])
    t = ol.OracleTraj(s, w["q"], w["p"], w["beta"], ensemble=0)
    for block in range(4):
        e.run(25)
        for _ in range(25):
            t.step()
        q, p = e.get_state()
        err = np.abs(q[0] - t.q).max() / np.abs(t.q).max()
        perr = np.abs(p[0] - t.p).max() / np.abs(t.p).max()
        print(f"O3 nb={nb} nm_kernel={nmk} step={25 * (block + 1)}: rel dq {err:.1e} dp {perr:.1e}")
        assert err <= 1e-10 and perr <= 1e-9
    e.close()


def _o3_children(ntraj, nb, seed):
    w = workloads.o3_nve(nbeads=nb)
    rng = np.random.default_rng(seed)
    q = workloads.O3_TS[None, None] + rng.normal(0.0, 0.02, (ntraj, 1, 3, 3))
    q = q + rng.normal(0.0, 0.005, (ntraj, nb, 3, 3))
    return w, q, np.full(ntraj, w["beta"])


def test_o3_children_and_kappa_match_oracle():
    ntraj, nb, nsteps = 12, 8, 60
    w, q, beta = _o3_children(ntraj, nb, 3)
    rx = workloads.o3_reaction()
    e = cam.RPMDEngine(cam.CAM_SURFACE_O3, nb, ntraj, w["dt"], w["masses"], ensemble=cam.CAM_ENS_CHILD,
                       reaction=rx, seed=5, steps_per_launch=25)
    e.set_state(q, np.zeros_like(q), beta)
    assert not e.constrain(project_p=False).any()
    e.draw_momenta(draw=1, antithetic_pairs=True)
    q0, p0 = e.get_state()
    e.begin_children()
    e.run(nsteps)
    num, den = e.kappa(nsteps)
    q1, _ = e.get_state()
    s = ol.make_sys(5, nb, w["dt"], w["masses"])
    orx = ol.make_rx(rx)
    num_ref, den_ref = np.zeros(nsteps), 0.0
    for t in range(ntraj):
        wt = ol.rpmd().oracle_child_weight(ol.ctypes.byref(s), ol.ctypes.byref(orx),
                                           np.ascontiguousarray(q0[t]).ctypes.data_as(ol._dp),
                                           np.ascontiguousarray(p0[t]).ctypes.data_as(ol._dp), float(beta[t]))
        if wt > 0:
            den_ref += wt
        tr = ol.OracleTraj(s, q0[t], p0[t], float(beta[t]), rx=orx, ensemble=2)
        for step in range(nsteps):
            if tr.step() > 0:
                num_ref[step] += wt
        np.testing.assert_allclose(q1[t], tr.q, rtol=0, atol=1e-9)
    print(f"O3 children: kappa(t_end) {num[-1] / den:.4f}, oracle {num_ref[-1] / den_ref:.4f}")
    assert den > 0
    np.testing.assert_allclose(den, den_ref, rtol=1e-12)
    np.testing.assert_allclose(num, num_ref, rtol=1e-12, atol=1e-12)
    e.close()


def test_o3_umbrella_window_matches_oracle():
    nwin, nb, ast, seed, n_equi, n_samp = 4, 8, 5, 11, 15, 25
    w, q, beta = _o3_children(nwin, nb, 7)
    rng = np.random.default_rng(8)
    p = rng.normal(size=q.shape) * np.sqrt(w["masses"][None, None, :, None] * nb / beta[:, None, None, None])
    rx = workloads.o3_reaction()
    xi_w, k_w = np.linspace(0.95, 1.05, nwin), np.full(nwin, 0.1)
    e = cam.RPMDEngine(cam.CAM_SURFACE_O3, nb, nwin, w["dt"], w["masses"], ensemble=cam.CAM_ENS_UMBRELLA,
                       reaction=rx, andersen_step=ast, seed=seed)
    e.set_state(q, p, beta)
    e.set_windows(xi_w, k_w)
    e.begin_umbrella(bias=True)
    s = ol.make_sys(5, nb, w["dt"], w["masses"])
    refs = []
    for t in range(nwin):
        orx = ol.make_rx(rx)
        orx.xi_ideal = xi_w[t]
        tr = ol.OracleTraj(s, q[t], p[t], float(beta[t]), rx=orx, ensemble=4, andersen_step=ast, seed=seed, traj=t)
        tr.umbrella_begin(k_w[t], bias=True)
        refs.append(tr)
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
    qe, _ = e.get_state()
    avg, var = e.umbrella_stats()
    for t, tr in enumerate(refs):
        err = np.abs(qe[t] - tr.q).max() / np.abs(tr.q).max()
        a_ref = tr.xi_sum / n_samp
        print(f"O3 window {xi_w[t]:.3f}: rel dq {err:.1e}; <xi> {avg[t]:.6f} (oracle {a_ref:.6f})")
        assert err <= 1e-10
        assert abs(avg[t] - a_ref) <= 1e-12
        assert abs(var[t] - (tr.xi_sq / n_samp - a_ref * a_ref)) <= 1e-12
    e.close()


def test_o3_engine_rejects_mixed_precision():
    w = workloads.o3_nve(nbeads=4)
    with pytest.raises(cam.CamError, match="FP64 only"):
        cam.RPMDEngine(cam.CAM_SURFACE_O3, 4, 1, w["dt"], w["masses"], precision=cam.CAM_PREC_MIXED)
