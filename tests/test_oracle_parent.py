"""The oracle's recrossing parent (verlet with constrain = 1): pinned by what
the constraint defines rather than by retyping its formulas.  After every
step the centroid lies on the dividing surface (mode-2 xi = 0) and, except on
Andersen steps, the bead momenta have no component along the constraint
direction of the step; Andersen steps draw the momenta the counter-based
stream gives; the start direction is the gradient of the mode-1 coordinate
(central differences)."""
import numpy as np

import oracle_lib as ol
from caracal_amd import workloads


def _parent(nb=8, seed=3, andersen_step=7):
    w = workloads.ch4h_children(ntraj=1, nbeads=nb, seed=seed)
    s = ol.make_sys(2, nb, w["dt"], w["masses"])
    rx = ol.make_rx(w["reaction"])
    rng = np.random.default_rng(seed)
    p = rng.normal(size=w["q"][0].shape) * np.sqrt(w["masses"][None, :, None] * nb / w["beta"][0])
    t = ol.OracleTraj(s, w["q"][0], p, float(w["beta"][0]), rx=rx, ensemble=3, andersen_step=andersen_step,
                      seed=11, traj=2)
    return w, s, rx, t


def test_parent_stays_on_the_dividing_surface_and_projects_momenta():
    w, s, rx, t = _parent()
    t.parent_begin()
    masses = w["masses"]
    worst = 0.0
    for i in range(1, 41):
        assert t.step() == 0
        xi, _ = ol.calc_xi(s, rx, t.q.mean(axis=0), 2)
        # constrain_q's Newton iteration stops once |xi| < 1e-10 (single-precision literal)
        worst = max(worst, abs(xi))
        assert abs(xi) < 1.5e-10, (i, xi)
        if i % t.andersen_step != 0:
            flux = np.sum(t.dxi[None] * t.p / masses[None, :, None])
            scale = np.sum(np.abs(t.dxi[None] * t.p / masses[None, :, None]))
            assert abs(flux) < 1e-13 * scale, (i, flux)
    print(f"largest |xi| after a step: {worst:.1e}")


def test_parent_andersen_steps_draw_the_counter_stream():
    w, s, rx, t = _parent(nb=4, andersen_step=5)
    t.parent_begin()
    for _ in range(5):
        t.step()
    nb, na = t.q.shape[0], t.q.shape[1]
    sd = np.sqrt(w["masses"] / (t.beta / nb))
    for k in range(nb):
        for a in range(na):
            for d in range(3):
                ref = ol.andersen_normal(11, 2, 5, 3 * a + d, k) * sd[a]
                assert t.p[k, a, d] == ref


def test_parent_begin_direction_is_the_mode1_gradient():
    w, s, rx, t = _parent()
    t.parent_begin()
    c = t.q.mean(axis=0)
    h = 1e-6
    fd = np.zeros_like(c)
    for a in range(c.shape[0]):
        for d in range(3):
            cp, cm = c.copy(), c.copy()
            cp[a, d] += h
            cm[a, d] -= h
            fd[a, d] = (ol.calc_xi(s, rx, cp, 1)[0] - ol.calc_xi(s, rx, cm, 1)[0]) / (2 * h)
    np.testing.assert_allclose(t.dxi, fd, rtol=0, atol=1e-7 * np.abs(fd).max())


def _off_surface_centroid(seed=3):
    w, s, rx, t = _parent(seed=seed)
    c = t.q.mean(axis=0) + np.random.default_rng(seed).normal(0, 0.05, t.q.shape[1:])
    return w, s, rx, c


def test_xi_hessian_is_the_jacobian_of_the_gradient():
    """calc_xi's Hessian (both modes) against central differences of the
    gradient it returns, and symmetric."""
    w, s, rx, c = _off_surface_centroid()
    h = 1e-5
    for mode in (1, 2):
        _, _, H = ol.calc_xi_hess(s, rx, c, mode)
        fd = np.zeros_like(H)
        flat = c.reshape(-1)
        for b in range(flat.size):
            cp, cm = flat.copy(), flat.copy()
            cp[b] += h
            cm[b] -= h
            gp = ol.calc_xi(s, rx, cp.reshape(c.shape), mode)[1].reshape(-1)
            gm = ol.calc_xi(s, rx, cm.reshape(c.shape), mode)[1].reshape(-1)
            fd[:, b] = (gp - gm) / (2 * h)
        scale = np.abs(H).max()
        np.testing.assert_allclose(H, fd, rtol=0, atol=1e-7 * scale)
        np.testing.assert_allclose(H, H.T, rtol=0, atol=1e-12 * scale)


def test_umbrella_bias_gradient_is_the_gradient_of_its_energy():
    """umbrella mode 0: the gradient added to every bead (force-constant term
    plus the Hessian metric term) is the derivative of the bias energy
    0.5 k (xi - xi_ideal)^2 - log(fs)/beta, by central differences; with k = 0
    only the metric term is left, and it is not zero."""
    w, s, rx, c = _off_surface_centroid(seed=5)
    rx.xi_ideal = 0.95                  # a window away from this centroid's xi
    beta = float(w["beta"][0])
    for k in (0.0, 0.37):
        v0, xi, dxi, g = ol.umbrella(s, rx, c, beta, k)
        assert np.array_equal(g[0], g[-1])
        fd = np.zeros_like(c)
        h = 1e-6
        for a in range(c.shape[0]):
            for d in range(3):
                cp, cm = c.copy(), c.copy()
                cp[a, d] += h
                cm[a, d] -= h
                fd[a, d] = (ol.umbrella(s, rx, cp, beta, k)[0] - ol.umbrella(s, rx, cm, beta, k)[0]) / (2 * h)
        np.testing.assert_allclose(g[0], fd, rtol=0, atol=1e-6 * np.abs(fd).max())
        assert np.abs(g[0]).max() > 0
        if k:
            assert abs(xi - rx.xi_ideal) > 1e-2
