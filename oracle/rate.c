/* Oracle (test infrastructure only): the rate-theory evaluation of
 * calc_rate after the umbrella samplings -- umbrella integration of the
 * window statistics into the PMF, its extrema, and the bimolecular QTST rate.
 * Restated from src/calc_rate.f90:1942-2092 and src/calc_k_t.f90:49-146. */
#include <math.h>
#include <stdlib.h>
#include "oracle.h"

static const double kPi = 3.1415926535897932384626433832795029; /* qmdff module constant */

/* Umbrella integration (Kaestner and Thiel; src/calc_rate.f90:1942-1999)
 * over nwin windows: window i has centre xi_wins[i], force constant
 * k_force[i] and the mean and variance of xi.  Writes dA[nbins] (the mean
 * force at bin j, xi = xi_min + (j) * bin_size for j = 0..nbins-1),
 * bin_coord[nbins-1] and pmf[nbins]: the trapezoid integral of dA at
 * bin_coord[i] for i < nbins-1, pmf[nbins-1] left at 0 as the reference
 * leaves it, and everything shifted by the minimum over all nbins entries. */
void oracle_umbrella_integration(int nwin, const double *xi_wins, const double *average, const double *variance,
                                 const double *k_force, double beta, double xi_min, double xi_max, int nbins,
                                 int umbr_traj, int umbr_step, double *dA, double *bin_coord, double *pmf)
{
    const double bin_size = (xi_max - xi_min) / (nbins - 1);
    double *p_ib = malloc(sizeof(double) * (size_t)nwin), *dA_iu = malloc(sizeof(double) * (size_t)nwin);
    for (int j = 0; j < nbins; ++j) {
        double denom = 0.0;
        const double xi_act = xi_min + bin_size * j;
        dA[j] = 0.0;
        for (int i = 0; i < nwin; ++i) {
            p_ib[i] = 1.0 / (sqrt(2.0 * kPi * variance[i])) *
                      exp(-0.5 * ((xi_act - average[i])) * ((xi_act - average[i])) / variance[i]);
            dA_iu[i] = (1.0 / beta) * (xi_act - average[i]) / (variance[i]) - k_force[i] * (xi_act - xi_wins[i]);
        }
        for (int k = 0; k < nwin; ++k) denom = denom + (double)(umbr_traj * umbr_step) * p_ib[k];
        for (int k = 0; k < nwin; ++k) dA[j] = dA[j] + (double)(umbr_traj * umbr_step) * p_ib[k] * dA_iu[k];
        dA[j] = dA[j] / denom;
    }
    free(p_ib);
    free(dA_iu);
    double A_act = 0.0;
    const double dx = bin_size;
    for (int i = 0; i < nbins; ++i) pmf[i] = 0.0;
    for (int i = 0; i < nbins - 1; ++i) {
        bin_coord[i] = 0.5 * (xi_min + i * dx + xi_min + (i + 1) * dx);
        A_act = A_act + 0.5 * dx * (dA[i] + dA[i + 1]);
        pmf[i] = A_act;
    }
    double mn = pmf[0];
    for (int i = 1; i < nbins; ++i)
        if (pmf[i] < mn) mn = pmf[i];
    for (int i = 0; i < nbins; ++i) pmf[i] = pmf[i] - mn;
}

/* The PMF's maximum and the reference point of the rate formula
 * (src/calc_rate.f90:2059-2092): maxloc over all nbins entries (first
 * maximum), and either the bin of xi = 0 (minloc "ZERO",
 * -int(xi_min/((xi_max-xi_min)/nbins)) + 1, 1-based) or the first minimum
 * left of the maximum.  Returns 0-based indices. */
void oracle_pmf_extrema(const double *pmf, int nbins, double xi_min, double xi_max, int minloc_zero, int *imax,
                        int *imin)
{
    int mx = 0;
    for (int i = 1; i < nbins; ++i)
        if (pmf[i] > pmf[mx]) mx = i;
    *imax = mx;
    if (minloc_zero) {
        *imin = -(int)(xi_min / ((xi_max - xi_min) / (nbins))) + 1 - 1;
    } else {
        int mn = 0;
        for (int i = 1; i <= mx; ++i)
            if (pmf[i] < pmf[mn]) mn = i;
        *imin = mn;
    }
}

/* Bimolecular QTST rate (src/calc_k_t.f90:126-129) in cm^3/(mol s):
 * n_paths kappa 4 pi R_inf^2 sqrt(1/(2 pi beta mu)) exp(-beta (W_max - W_min)),
 * mu the reduced mass of the two reactants (electron masses), then the unit
 * factor 1e6 (a0^3 / t_au) N_A, evaluated as the reference evaluates its
 * default-REAL literals: in single precision. */
double oracle_rate_bimolecular(int npaths, double kappa, double r_inf, double beta, double m1, double m2,
                               double pmf_max, double pmf_min)
{
    const double pi = 3.1415926535897932384;
    const double my_R = m1 * m2 / (m1 + m2);
    double k_t = npaths * kappa * 4.0 * pi * r_inf * r_inf * sqrt(1.0 / (2.0 * pi * beta * my_R)) *
                 exp(-beta * (pmf_max - pmf_min));
    const float a0 = 5.2917721092e-11f, tau = 2.418884326505e-17f, na = 6.02214179e23f;
    const float a3 = a0 * a0 * a0;
    const float unit = a3 / tau;
    k_t = k_t * (double)1e6f * (double)unit * (double)na;
    return k_t;
}
