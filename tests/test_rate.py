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


def _wham_golden():
    import os
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "wham_reference.npz"))
    for n in range(int(z["ncases"])):
        yield {k[len(f"c{n}_"):]: z[k] for k in z.files if k.startswith(f"c{n}_")}


def test_wham_matches_the_compiled_reference():
    """rate.wham against src/wham.f90 compiled from the reference
    (tests/golden/make_wham_golden.py): multinomial window histograms of a
    double well, and a range twice the windows' span whose empty bins take
    the routine's 1e-15 floor.  The PMF in the reference's stored units."""
    for c in _wham_golden():
        coord, pmf, pref, it = rate.wham(c["hist"], c["xi_wins"], c["k_force"], float(c["xi_min"]),
                                         float(c["xi_max"]), int(c["nbins"]), float(c["kelvin"]))
        ref = c["pmf_ref"]
        assert np.all(np.isfinite(ref))
        np.testing.assert_allclose(pref, ref, rtol=0, atol=1e-13 * np.abs(ref).max())
        np.testing.assert_allclose(pmf * rate.KCAL_PER_HARTREE * rate.BOHR_ANGSTROM, pref, rtol=1e-15)
        assert pmf.min() == 0.0
        bs = (float(c["xi_max"]) - float(c["xi_min"])) / (int(c["nbins"]) - 1)
        np.testing.assert_allclose(np.diff(coord), bs, rtol=1e-12)


def test_wham_recovers_the_pmf_from_exact_histograms():
    """Histograms proportional to exp(-beta (A + V_i)) on the bin
    coordinates are WHAM's fixed point: converged tightly, the PMF is A up to
    a constant (a wrong bias coordinate, sign or normalisation moves it)."""
    kelvin = 300.0
    beta = 1.0 / rate.KB_HARTREE / kelvin
    xi_min, xi_max, nbins = -0.2, 1.2, 281
    bs = (xi_max - xi_min) / (nbins - 1)
    x = xi_min + bs * np.arange(1, nbins + 1)
    A = 0.004 * np.sin(4.0 * x) + 0.002 * x
    xw, k = np.linspace(-0.1, 1.1, 13), np.linspace(0.04, 0.07, 13)
    h = np.array([np.exp(-beta * (A + ki * (x - xi) ** 2)) for xi, ki in zip(xw, k)])
    h = 1e6 * h / h.sum(axis=1)[:, None]
    _, pmf, _, _ = rate.wham(h, xw, k, xi_min, xi_max, nbins, kelvin, tol=1e-20)
    np.testing.assert_allclose(pmf, A - A.min(), rtol=0, atol=1e-10)
    # the default tolerance stops early but close
    _, pmf5, _, it = rate.wham(h, xw, k, xi_min, xi_max, nbins, kelvin)
    assert it < 100 and np.abs(pmf5 - (A - A.min())).max() < 1e-4


def test_wham_needs_samples_in_every_window():
    h = np.ones((3, 11))
    h[1] = 0.0
    try:
        rate.wham(h, [0.0, 0.5, 1.0], 0.05, 0.0, 1.0, 11, 300.0)
    except ValueError:
        return
    raise AssertionError("an empty window must be rejected")
