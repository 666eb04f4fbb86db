"""Rate theory after the umbrella samplings (caracal_amd.rate) against the
oracle's restatement of calc_rate / calc_k_t, and the oracle pinned by what
the mathematics fixes: umbrella integration is exact for the Gaussian
windows of a harmonic PMF, and the rate's unit factor is the atomic unit of
a bimolecular rate coefficient."""
import numpy as np

import oracle_lib as ol
from caracal_amd import rate


def _harmonic_windows(a=0.03, c=0.55, beta=1052.6, nwin=23):
    """Windows on A(xi) = a/2 (xi - c)^2: the biased distribution of window
    i is Gaussian with variance 1/(beta (a + k_i)) and mean
    (a c + k_i xi_i)/(a + k_i)."""
    xi_w = np.linspace(-0.05, 1.05, nwin)
    k_w = np.linspace(0.04, 0.08, nwin)
    var = 1.0 / (beta * (a + k_w))
    avg = (a * c + k_w * xi_w) / (a + k_w)
    return xi_w, avg, var, k_w, beta, a, c


def test_oracle_umbrella_integration_is_exact_for_a_harmonic_pmf():
    xi_w, avg, var, k_w, beta, a, c = _harmonic_windows()
    xi_min, xi_max, nbins = -0.05, 1.05, 501
    dA, bc, pmf = ol.umbrella_integration(xi_w, avg, var, k_w, beta, xi_min, xi_max, nbins, 10, 20000)
    grid = xi_min + (xi_max - xi_min) / (nbins - 1) * np.arange(nbins)
    np.testing.assert_allclose(dA, a * (grid - c), rtol=0, atol=1e-12)
    # pmf[i], reported at the centre of bin i, is the integral up to its right edge
    A = 0.5 * a * (grid - c) ** 2
    raw = np.append(A[1:] - A[0], 0.0)
    np.testing.assert_allclose(pmf, raw - raw.min(), rtol=0, atol=1e-12)
    np.testing.assert_allclose(bc, 0.5 * (grid[:-1] + grid[1:]), rtol=0, atol=1e-15)


def test_umbrella_integration_matches_oracle():
    rng = np.random.default_rng(3)
    nwin = 40
    xi_w = np.linspace(-0.05, 1.05, nwin)
    avg = xi_w + rng.normal(0, 0.01, nwin)
    var = rng.uniform(2e-5, 4e-4, nwin)
    k_w = rng.uniform(0.03, 0.07, nwin)
    for nbins in (101, 5000):
        ref = ol.umbrella_integration(xi_w, avg, var, k_w, 950.0, -0.05, 1.05, nbins, 10, 20000)
        got = rate.umbrella_integration(xi_w, avg, var, k_w, 950.0, -0.05, 1.05, nbins, 10, 20000)
        for r, g in zip(ref, got):
            np.testing.assert_allclose(g, r, rtol=1e-11, atol=1e-13 * np.abs(r).max())


def test_pmf_extrema_match_oracle():
    xi_min, xi_max, nbins = -0.05, 1.05, 5000
    x = np.linspace(xi_min, xi_max, nbins)
    pmf = 0.02 * np.exp(-((x - 0.93) / 0.1) ** 2) - 0.003 * np.exp(-((x - 0.2) / 0.05) ** 2)
    for zero in (True, False):
        ref = ol.pmf_extrema(pmf, xi_min, xi_max, zero)
        got = rate.pmf_extrema(pmf, xi_min, xi_max, "ZERO" if zero else "PMF_MIN")
        assert ref == got
    imax, imin = rate.pmf_extrema(pmf, xi_min, xi_max, "ZERO")
    assert abs(x[imax] - 0.93) < 1e-3
    # the ZERO reference point is the bin at xi = 0 (within a bin)
    assert abs(xi_min + (imin + 0.5) * (xi_max - xi_min) / (nbins - 1)) < 2 * (xi_max - xi_min) / nbins
    assert abs(x[rate.pmf_extrema(pmf, xi_min, xi_max, "PMF_MIN")[1]] - 0.2) < 1e-3


def test_rate_matches_oracle_and_the_atomic_unit_of_a_rate_coefficient():
    args = (2, 0.61, 16.0, 1052.58, 1837.15 * 2.016, 1837.15 * 1.008, 0.0161, 0.0002)
    ref = ol.rate_bimolecular(*args)
    got = rate.rate_bimolecular(args[1], args[3], args[6], args[7], (args[4], args[5]), args[2], n_paths=args[0])
    assert got == ref
    # with every physical factor 1: 1e6 a0^3/t_au N_A, a0 = 5.29177e-11 m, t_au = 2.41888e-17 s
    unit = rate.rate_bimolecular(1.0, 1.0 / (2 * np.pi), 0.0, 0.0, (2.0, 2.0), 1.0 / np.sqrt(4 * np.pi))
    assert abs(unit / (1e6 * 5.29177210903e-11 ** 3 / 2.4188843265857e-17 * 6.02214076e23) - 1) < 1e-6
    # proportional to kappa and to exp(-beta dW), and to 1/sqrt(reduced mass)
    k1 = rate.rate_bimolecular(0.5, 1000.0, 0.01, 0.0, (4.0, 4.0), 16.0)
    assert abs(rate.rate_bimolecular(1.0, 1000.0, 0.01, 0.0, (4.0, 4.0), 16.0) / k1 - 2) < 1e-14
    assert abs(rate.rate_bimolecular(0.5, 1000.0, 0.011, 0.0, (4.0, 4.0), 16.0) / k1 - np.exp(-1.0)) < 1e-12
    assert abs(rate.rate_bimolecular(0.5, 1000.0, 0.01, 0.0, (16.0, 16.0), 16.0) / k1 - 0.5) < 1e-14
