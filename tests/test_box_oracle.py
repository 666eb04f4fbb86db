"""The numpy ring-polymer step of the box oracle (tests/box_oracle.py) is
pinned to the C oracle's free ring-polymer step (oracle_free_rp, itself pinned
by its closed form) for the reference transform, and to the exact
conservation of the free ring-polymer energy for the orthogonal basis; its
Nose-Hoover chain to the C oracle's (itself pinned by the extended-system
energy)."""
import numpy as np

import box_oracle as bo
import oracle_lib as ol
from caracal_amd import workloads


def test_numpy_free_rp_matches_c_oracle_reference_mode():
    rng = np.random.default_rng(0)
    for nb in (2, 4, 8, 16, 32):
        masses = rng.uniform(1000, 30000, 5)
        q = rng.normal(size=(nb, 5, 3))
        p = rng.normal(size=(nb, 5, 3)) * 30
        dt, beta = 4.13, 1000.0
        s = ol.make_sys(1, nb, dt, masses)
        o = ol.rpmd()
        qc, pc = q.copy(), p.copy()
        o.oracle_free_rp(s, ol._ptr(qc), ol._ptr(pc), beta)
        qn, pn = bo.free_rp(q, p, dt, masses, beta)
        np.testing.assert_allclose(qn, qc, rtol=0, atol=1e-13 * np.abs(qc).max())
        np.testing.assert_allclose(pn, pc, rtol=0, atol=1e-13 * np.abs(pc).max())


def test_orthogonal_free_rp_conserves_the_free_ring_polymer_energy():
    rng = np.random.default_rng(1)
    for nb in (4, 7, 16):
        masses = rng.uniform(1000, 30000, 4)
        q = rng.normal(size=(nb, 4, 3))
        p = rng.normal(size=(nb, 4, 3)) * 30
        dt, beta = 4.13, 800.0
        wn = nb / beta

        def h(q, p):
            d = q - np.roll(q, -1, axis=0)
            return 0.5 * np.sum(p ** 2 / masses[None, :, None]) + 0.5 * wn * wn * np.sum(masses[None, :, None] * d * d)
        h0 = h(q, p)
        for _ in range(50):
            q, p = bo.free_rp(q, p, dt, masses, beta, orthogonal=True)
        assert abs(h(q, p) - h0) < 1e-10 * h0
        u = bo.orthogonal_matrix(nb)
        np.testing.assert_allclose(u @ u.T, np.eye(nb), atol=1e-14)


def test_numpy_nhc_matches_c_oracle():
    """The box oracle's chain (numpy) against the C oracle's oracle_nhc on a
    three-atom ring polymer, over 30 successive applications."""
    import box_oracle as bo
    w = workloads.h3_nve(nbeads=8)
    s = ol.make_sys(1, 8, w["dt"], w["masses"])
    p_c = np.ascontiguousarray(w["p"] * 1.4)
    p_n = p_c.copy()
    v_c, v_n = np.zeros(4), [0.0] * 4
    nose_q = 15.0 * w["dt"]
    kelvin = 1.0 / (w["beta"] * 0.316679e-5)
    qterm = 0.316679e-5 * kelvin * nose_q * nose_q
    qnh = np.array([9 * qterm, qterm, qterm, qterm])
    for _ in range(30):
        ol.rpmd().oracle_nhc(ol.ctypes.byref(s), ol._ptr(p_c), ol._ptr(v_c), ol._ptr(qnh), kelvin, None)
        p_n = bo.nhc(p_n, np.asarray(w["masses"]), v_n, nose_q, w["beta"], w["dt"])
    np.testing.assert_allclose(p_n, p_c, rtol=1e-12, atol=0)
    np.testing.assert_allclose(v_n, v_c, rtol=1e-12, atol=1e-30)
    assert abs(v_c[0]) > 0


def test_numpy_transrot_equals_the_c_oracle():
    """box_oracle.transrot (any size) against oracle_transrot (up to 16 atoms)."""
    import ctypes
    import oracle_lib as ol
    from box_oracle import transrot
    rng = np.random.default_rng(8)
    for nb, na in ((1, 5), (4, 7), (8, 16)):
        m = rng.uniform(1000, 30000, na)
        q = rng.normal(0, 3, (nb, na, 3)) + 10.0
        p = rng.normal(0, 5, (nb, na, 3))
        s = ol.make_sys(1, nb, 1.0, m)
        pc = np.ascontiguousarray(p.copy())
        ol.rpmd().oracle_transrot(ctypes.byref(s), np.ascontiguousarray(q).ctypes.data_as(ol._dp),
                                  pc.ctypes.data_as(ol._dp))
        assert np.abs(transrot(q, p, m) - pc).max() <= 1e-11 * np.abs(pc).max()
