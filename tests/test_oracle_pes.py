"""Pins the oracle's analytic surfaces: against the compiled reference
(oracle/_ref), against central differences of its own energy (the check the
reference's gradnum/NUM_GRAD performs), and against symmetries of the surface."""
import numpy as np
import pytest

import geometries as geo
import oracle_lib as ol

needs_ref = pytest.mark.skipif(ol.ref() is None, reason="compiled reference not built")


def _fd(fn, q, h=1e-5):
    g = np.zeros_like(q)
    for b in range(q.shape[0]):
        for i in range(q.shape[1]):
            for d in range(3):
                qp = q.copy(); qp[b, i, d] += h
                qm = q.copy(); qm[b, i, d] -= h
                g[b, i, d] = (fn(qp)[0][b] - fn(qm)[0][b]) / (2 * h)
    return g


@needs_ref
@pytest.mark.parametrize("q", [geo.h3_cloud(300, 0.3, 0), geo.h3_special()], ids=["cloud", "special"])
def test_h3_matches_compiled_reference(q):
    v, g = ol.egrad_h3(q)
    vr, gr = ol.ref_egrad_h3(q)
    np.testing.assert_allclose(v, vr, rtol=1e-13, atol=0)
    np.testing.assert_allclose(g, gr, rtol=0, atol=1e-13 * np.abs(gr).max())


@needs_ref
@pytest.mark.parametrize("dhb", [1.4, 2.0, 5.0, 20.0])
def test_ch4h_matches_compiled_reference(dhb):
    q = geo.ch4h_cloud(200, 0.1, 1, d_hb=dhb)
    v, g = ol.egrad_ch4h(q)
    vr, gr = ol.ref_egrad_ch4h(q)
    np.testing.assert_allclose(v, vr, rtol=1e-13, atol=0)
    np.testing.assert_allclose(g, gr, rtol=0, atol=1e-13 * np.abs(gr).max())


def test_h3_gradient_is_derivative_of_energy():
    # asymmetric members only: |(R1-R2)(R2-R3)(R3-R1)| has a kink where two
    # distances coincide, so a central difference straddles it there
    q = np.concatenate([geo.h3_cloud(4, 0.2, 3), geo.h3_special()[[2, 5, 8, 14]]])
    _, g = ol.egrad_h3(q)
    np.testing.assert_allclose(_fd(ol.egrad_h3, q), g, atol=2e-8)


def test_ch4h_gradient_is_derivative_of_energy():
    q = geo.ch4h_cloud(3, 0.08, 4)
    _, g = ol.egrad_ch4h(q)
    np.testing.assert_allclose(_fd(ol.egrad_ch4h, q), g, atol=5e-7)


def test_h3_permutation_and_rigid_motion_invariance():
    q = geo.h3_cloud(50, 0.3, 5)
    v, g = ol.egrad_h3(q)
    for perm in ([1, 0, 2], [2, 1, 0], [0, 2, 1], [1, 2, 0]):
        vp, gp = ol.egrad_h3(q[:, perm])
        np.testing.assert_allclose(vp, v, rtol=1e-13)
        np.testing.assert_allclose(gp, g[:, perm], atol=1e-12)
    rot = np.linalg.qr(np.random.default_rng(6).normal(size=(3, 3)))[0]
    vr, gr = ol.egrad_h3(q @ rot.T + 3.7)
    np.testing.assert_allclose(vr, v, rtol=1e-12)
    np.testing.assert_allclose(gr, g @ rot.T, atol=1e-11)
    # forces on an isolated system sum to zero
    np.testing.assert_allclose(g.sum(axis=1), 0.0, atol=1e-12)


def test_h3_collinear_saddle_is_stationary_along_symmetric_stretch():
    # At the collinear saddle R1 = R2 the gradient along the antisymmetric
    # stretch vanishes and the surface is symmetric under swapping the ends.
    v, g = ol.egrad_h3(geo.H3_TS[None])
    assert abs(g[0, 0, 2] + g[0, 2, 2]) < 1e-12
    assert abs(g[0, 1, 2]) < 1e-12
    # classical barrier of BKMP2: about 9.6-9.7 kcal/mol above H + H2(re)
    r = np.array([[0, 0, 0], [1.401, 0, 0], [40.0, 0, 0]])
    v_asym, _ = ol.egrad_h3(r[None])
    barrier = (v[0] - v_asym[0]) * 627.509
    assert 9.4 < barrier < 9.9


def test_ch4h_rotation_invariance_and_zero_net_force():
    q = geo.ch4h_cloud(20, 0.1, 7)
    v, g = ol.egrad_ch4h(q)
    rot = np.linalg.qr(np.random.default_rng(8).normal(size=(3, 3)))[0]
    if np.linalg.det(rot) < 0:
        rot[:, 0] *= -1
    vr, gr = ol.egrad_ch4h(q @ rot.T - 1.3)
    np.testing.assert_allclose(vr, v, rtol=1e-11)
    np.testing.assert_allclose(gr, g @ rot.T, atol=1e-10)
    np.testing.assert_allclose(g.sum(axis=1), 0.0, atol=1e-10)
