"""Rate theory after the umbrella samplings, as calc_rate evaluates it:
umbrella integration of the window statistics into the potential of mean
force (src/calc_rate.f90:1942-1999), the PMF's extrema
(src/calc_rate.f90:2059-2092), and the bimolecular QTST rate constant
(src/calc_k_t.f90:49-146).  Host arithmetic in float64; the window
statistics come from the engine (ensemble UMBRELLA) and kappa from
caracal_amd.recross."""
import numpy as np

PI = 3.1415926535897932384626433832795029
AVOGADRO = 6.02214179e+23


def umbrella_integration(xi_wins, average, variance, k_force, beta, xi_min, xi_max, nbins, umbr_traj=1, umbr_step=1):
    """Umbrella integration (Kaestner and Thiel) over the windows given.

    Returns (dA, bin_coord, pmf): the mean force at the nbins grid points
    xi_min + j (xi_max - xi_min)/(nbins - 1); the nbins - 1 bin centres; and
    the PMF at those centres (hartree), the trapezoid integral of dA.  As in
    the reference, the PMF is shifted by its minimum over nbins entries of
    which the last, never assigned, is 0; pmf has nbins entries, the last
    being that unassigned one."""
    xi_wins, average, variance, k_force = (np.asarray(a, dtype=np.float64) for a in (xi_wins, average, variance,
                                                                                      k_force))
    bin_size = (xi_max - xi_min) / (nbins - 1)
    xi_act = xi_min + bin_size * np.arange(nbins, dtype=np.float64)
    d = xi_act[:, None] - average[None, :]
    p_ib = 1.0 / np.sqrt(2.0 * PI * variance)[None, :] * np.exp(-0.5 * d * d / variance[None, :])
    dA_iu = (1.0 / beta) * d / variance[None, :] - k_force[None, :] * (xi_act[:, None] - xi_wins[None, :])
    w = float(umbr_traj * umbr_step) * p_ib
    dA = np.sum(w * dA_iu, axis=1) / np.sum(w, axis=1)
    dx = bin_size
    i = np.arange(nbins - 1, dtype=np.float64)
    bin_coord = 0.5 * (xi_min + i * dx + xi_min + (i + 1) * dx)
    pmf = np.zeros(nbins)
    pmf[:-1] = np.cumsum(0.5 * dx * (dA[:-1] + dA[1:]))
    return dA, bin_coord, pmf - pmf.min()


def pmf_extrema(pmf, xi_min, xi_max, minloc="ZERO"):
    """(imax, imin), 0-based: the first maximum of the PMF, and the
    reference point of the rate formula -- the bin of xi = 0 ("ZERO", the
    default for bimolecular reactions) or the first minimum left of the
    maximum ("PMF_MIN")."""
    pmf = np.asarray(pmf)
    nbins = len(pmf)
    imax = int(np.argmax(pmf))
    if minloc == "ZERO":
        imin = -int(xi_min / ((xi_max - xi_min) / nbins)) + 1 - 1
    elif minloc == "PMF_MIN":
        imin = int(np.argmin(pmf[: imax + 1]))
    else:
        raise ValueError("minloc must be 'ZERO' or 'PMF_MIN'")
    return imax, imin


def rate_bimolecular(kappa, beta, pmf_max, pmf_min, mass_reac, r_inf, n_paths=1):
    """QTST rate of a bimolecular reaction in cm^3/(mol s):
    n_paths kappa 4 pi R_inf^2 sqrt(1/(2 pi beta mu)) exp(-beta (W_max - W_min)),
    mu the reduced mass of the two reactants (electron masses, the
    reference's mass_reac).  The unit factor 1e6 (a0^3/t_au) N_A is formed in
    binary32, as the reference's default-REAL literals are.  Divide by
    AVOGADRO for cm^3/(molecule s)."""
    m1, m2 = mass_reac
    my_r = m1 * m2 / (m1 + m2)
    pi = 3.1415926535897932384
    k_t = n_paths * kappa * 4.0 * pi * r_inf * r_inf * np.sqrt(1.0 / (2.0 * pi * beta * my_r)) * \
        np.exp(-beta * (pmf_max - pmf_min))
    a0, tau, na = np.float32(5.2917721092e-11), np.float32(2.418884326505e-17), np.float32(6.02214179e23)
    unit = np.float32(np.float32(np.float32(a0 * a0) * a0) / tau)
    return float(k_t * float(np.float32(1e6)) * float(unit) * float(na))
