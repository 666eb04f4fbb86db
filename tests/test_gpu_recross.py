"""The recrossing driver (caracal_amd.recross) on the GPU against the same
calculation restated with the oracle: parents, children, draws and kappa(t)
for a small CH4 + H case, and a larger run's plausibility."""
import numpy as np
import pytest

import caracal_amd as cam
import oracle_lib as ol
from caracal_amd import workloads
from caracal_amd.recross import Recrossing

pytestmark = pytest.mark.gpu

NB, NPAR, CP, EVOL, INTERV, TIMES, EQUI, AST, KF, SEED = 8, 2, 4, 20, 4, 3, 8, 3, 0.2, 40


def _oracle_recrossing(w, q0, beta):
    """recross_serial restated with the oracle, with the driver's engines'
    keys (seed + creation index: parents (andersen 3), children, parents
    (andersen 0), parents (andersen 1)) and draw ids (2**40 + 1, ... in call
    order)."""
    s = ol.make_sys(2, NB, w["dt"], w["masses"])
    orx = ol.make_rx(w["reaction"])
    draw = [1 << 40]
    keys = {("parent", AST): SEED, "child": SEED + 1, ("parent", 0): SEED + 2, ("parent", 1): SEED + 3}

    def next_draw():
        draw[0] += 1
        return draw[0]

    def segment(qs, nsteps, ast):
        d = next_draw()
        out = []
        for t in range(NPAR):
            tr = ol.OracleTraj(s, qs[t], np.zeros_like(qs[t]), float(beta[t]), rx=orx, ensemble=3, andersen_step=ast,
                               seed=keys[("parent", ast)], traj=t)
            ol.rpmd().oracle_andersen(ol.ctypes.byref(s), ol._ptr(tr.p), float(beta[t]), keys[("parent", ast)], t, d)
            tr.parent_begin(KF)
            for _ in range(nsteps):
                assert tr.step() == 0
            out.append(tr.q.copy())
        return np.array(out)

    q = segment(q0, EQUI, AST)
    num, den = np.zeros(EVOL), 0.0
    for cycle in range(1, TIMES + 1):
        d = next_draw()
        for t in range(NPAR):
            for j in range(CP):
                ci = t * CP + j
                p = np.zeros_like(q[t])
                ol.rpmd().oracle_andersen(ol.ctypes.byref(s), ol._ptr(p), float(beta[t]), keys["child"], ci & ~1, d)
                if ci & 1:
                    p = -p
                wt = ol.rpmd().oracle_child_weight(ol.ctypes.byref(s), ol.ctypes.byref(orx), ol._ptr(q[t]), ol._ptr(p),
                                                   float(beta[t]))
                den += wt if wt > 0 else 0.0
                ch = ol.OracleTraj(s, q[t], p, float(beta[t]), rx=orx, ensemble=2)
                for l in range(EVOL):
                    if ch.step() > 0:
                        num[l] += wt
        q = segment(q, INTERV, 1 if cycle % 2 == 0 else 0)
    return q, num, den


def test_recrossing_driver_matches_oracle():
    w = workloads.ch4h_children(ntraj=NPAR, nbeads=NB, seed=6)
    q0, beta = w["q"], w["beta"] * np.array([1.0, 0.8])
    # parents start on the dividing surface, as after the reference's structure generation
    e = cam.RPMDEngine(cam.CAM_SURFACE_CH4H, NB, NPAR, w["dt"], w["masses"], ensemble=cam.CAM_ENS_CHILD,
                       reaction=w["reaction"])
    e.set_state(q0, np.zeros_like(q0), beta)
    assert not e.constrain(project_p=False).any()
    q0, _ = e.get_state()
    e.close()
    r = Recrossing(cam.CAM_SURFACE_CH4H, NB, w["dt"], w["masses"], w["reaction"], beta, child_point=CP,
                   child_evol=EVOL, child_interv=INTERV, child_times=TIMES, recr_equi=EQUI, andersen_step=AST,
                   k_force=KF, seed=SEED)
    out = r.run(q0)
    q_ref, num_ref, den_ref = _oracle_recrossing(w, q0, beta)
    err = np.abs(out["q_parents"] - q_ref).max() / np.abs(q_ref).max()
    print(f"parents rel dq {err:.1e}; kappa(t) {np.round(out['kappa_t'], 4)}")
    assert err <= 1e-10
    np.testing.assert_allclose(out["den"], den_ref, rtol=1e-10)
    np.testing.assert_allclose(out["num"], num_ref, rtol=0, atol=1e-10 * den_ref)
    assert out["errors"] == 0


def test_recrossing_kappa_is_plausible_for_many_parents():
    """64 parents x 64 children of CH4 + H at 500 K, 8 beads: kappa(t) starts
    near 1 (every child with positive flux is on the product side after one
    short step), stays within [0, 1.05] and decays; the parents stay on the
    dividing surface and pass the trajectory check."""
    nb, npar = 8, 64
    w = workloads.ch4h_children(ntraj=npar, nbeads=nb, seed=8)
    e = cam.RPMDEngine(cam.CAM_SURFACE_CH4H, nb, npar, w["dt"], w["masses"], ensemble=cam.CAM_ENS_CHILD,
                       reaction=w["reaction"])
    e.set_state(w["q"], np.zeros_like(w["q"]), w["beta"])
    assert not e.constrain(project_p=False).any()
    q0, _ = e.get_state()
    e.close()
    ts_energy = float(cam.egrad_ch4h(workloads.CH4H_TS[None])[0][0])
    r = Recrossing(cam.CAM_SURFACE_CH4H, nb, w["dt"], w["masses"], w["reaction"], w["beta"], child_point=64,
                   child_evol=400, child_interv=100, child_times=2, recr_equi=200, andersen_step=10,
                   check=dict(ts_energy=ts_energy, energy_tol=0.05, err_max=10), seed=3)
    out = r.run(q0)
    k = out["kappa_t"]
    print(f"kappa(t) at 1, 100, 200, 400 steps: {k[0]:.4f} {k[99]:.4f} {k[199]:.4f} {k[-1]:.4f}; "
          f"failed segments {out['errors']}")
    assert 0.9 <= k[0] <= 1.05
    assert (k >= 0).all() and (k <= 1.05).all()
    assert k[-1] <= k[0]
    s = ol.make_sys(2, nb, w["dt"], w["masses"])
    orx = ol.make_rx(w["reaction"])
    for t in range(npar):
        assert abs(ol.calc_xi(s, orx, out["q_parents"][t].mean(axis=0), 2)[0]) < 1.5e-10
