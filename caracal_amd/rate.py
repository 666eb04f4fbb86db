"""Rate theory after the umbrella samplings, as calc_rate evaluates it:
umbrella integration of the window statistics into the potential of mean
force (src/calc_rate.f90:1942-1999), or WHAM over the windows' histograms
of xi (src/wham.f90, called from src/calc_rate.f90:2006-2022), the PMF's
extrema (src/calc_rate.f90:2059-2092), and the bimolecular QTST rate
constant (src/calc_k_t.f90:49-146).  Host arithmetic in float64; the window
statistics and histograms come from the engine (ensemble UMBRELLA) and
kappa from caracal_amd.recross."""
import numpy as np

PI = 3.1415926535897932384626433832795029
AVOGADRO = 6.02214179e+23
KB_HARTREE = 0.316679e-5           # Boltzmann constant of src/wham.f90:82 (hartree/K)
KCAL_PER_HARTREE = 627.509         # FAC of src/wham.f90:86
BOHR_ANGSTROM = 0.52917721092      # bohr of src/general.f90:256


def _seqsum(a, axis):
    """Sum along axis in index order, one addition after the other (the
    order of the reference's DO loops; numpy's sum adds pairwise)."""
    return np.cumsum(a, axis=axis).take(-1, axis=axis)


def wham(hist, xi_wins, k_force, xi_min, xi_max, nbins, kelvin, tol=1.0e-5, max_iter=1000000):
    """WHAM (Roux 1995) as src/wham.f90 iterates it, over the windows given.

    hist: (nwin, nbins) sample counts per window, column j - 1 the bin of
    coordinate xi_min + j bin_size (bin_size = (xi_max - xi_min)/(nbins - 1),
    src/calc_rate.f90:1803), which is where the bias of window i,
    k_force_i (xi - xi_i)^2, is evaluated (src/wham.f90:104-106); kelvin the
    temperature (beta = 1/(0.316679e-5 kelvin), src/wham.f90:82-84).  From
    f_i = 1 it repeats P(j) = sum_i N_i p_i(j) / sum_i N_i w_i(j) / f_i (0
    becomes 1e-15) and f_i = sum_j w_i(j) P(j), w_i(j) = exp(-beta V_i(j)),
    until sum_i (1 - f_i new / f_i old)^2 <= tol (src/wham.f90:114-165).

    Returns (coord, pmf, pmf_ref_units, iterations): the bin coordinates
    xi_min + bin_size (j - 1/2) the reference writes (src/wham.f90:187), the
    PMF in hartree, -ln P(j)/beta shifted so that its minimum is 0
    (src/wham.f90:177-188), and the same times 627.509 x 0.52917721092, the
    units src/wham.f90:189 stores it in.

    Every window passed must hold samples.  The reference passes all n_all
    windows, of which the last never receives a histogram
    (src/calc_rate.f90:1808 reads n_all - 1 of them), and its normalisation
    of that empty one divides 0 by 0."""
    h = np.asarray(hist, dtype=np.float64)
    xi_wins = np.asarray(xi_wins, dtype=np.float64)
    k_force = np.broadcast_to(np.asarray(k_force, dtype=np.float64), xi_wins.shape)
    nw = len(xi_wins)
    if h.shape != (nw, nbins):
        raise ValueError(f"hist must be ({nw}, {nbins})")
    bin_size = (xi_max - xi_min) / (nbins - 1)
    beta = 1.0 / KB_HARTREE / float(kelvin)
    ni = _seqsum(h, axis=1)
    if np.any(ni <= 0.0):
        raise ValueError("every window needs samples")
    p_biased = h / ni[:, None]
    tmp2 = xi_min + bin_size * np.arange(1, nbins + 1, dtype=np.float64)
    d = tmp2[None, :] - xi_wins[:, None]
    v_res = k_force[:, None] * (d * d)
    w = np.exp(-beta * v_res)
    num = _seqsum(ni[:, None] * p_biased, axis=0)           # the same every iteration
    fi = np.ones(nw)
    for it in range(1, max_iter + 1):
        den = _seqsum((ni[:, None] * w) / fi[:, None], axis=0)
        p = num / den
        p = np.where(p == 0.0, 1.0e-15, p)
        fi_new = _seqsum(w * p[None, :], axis=1)
        eps = _seqsum((1.0 - fi_new / fi) ** 2, axis=0)
        fi = fi_new
        if not eps > tol:
            break
    else:
        raise RuntimeError(f"WHAM did not converge in {max_iter} iterations")
    lnp = (1.0 / beta) * np.log(p)
    pmin = max(-1.0e8, float(np.max(lnp)))
    pmf = -lnp + pmin
    coord = xi_min + bin_size * (np.arange(1, nbins + 1, dtype=np.float64) - 0.5)
    return coord, pmf, pmf * KCAL_PER_HARTREE * BOHR_ANGSTROM, it


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
