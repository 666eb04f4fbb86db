"""The numpy ring-polymer step of the box oracle (tests/box_oracle.py) is
pinned to the C oracle's free ring-polymer step (oracle_free_rp, itself pinned
by its closed form) for the reference transform, and to the exact
conservation of the free ring-polymer energy for the orthogonal basis."""
import numpy as np

import box_oracle as bo
import oracle_lib as ol


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
