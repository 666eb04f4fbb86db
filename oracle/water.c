/*
 * oracle/water.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * The reference's flexible water model, energy and gradient of one bead of a
 * box of nwater molecules in the order O, H, H (src/egrad_water.f90:36-333,
 * parameters src/water_init.f90:36-115), restated in plain C double:
 *   intramolecular  De (1 - exp(+a dr_OH))^2 for both O-H bonds (the
 *                   reference's exponent sign, :111-112), 0.5 k_HH dr_HH^2,
 *                   k_rt dr_HH (dr_OH1 + dr_OH2), k_rr dr_OH1 dr_OH2;
 *   intermolecular  q_i q_j / r, or the Zahn form q_i q_j (erfc(a r)/r -
 *                   zp (r - rc)) with the cutoff, for atoms of different
 *                   molecules, gradient e0/r^2 * d as the reference writes
 *                   it (:281-283, the derivative of q_i q_j / r);
 *                   LJ 4 eps ((s/r)^12 - (s/r)^6) whenever the FIRST atom of
 *                   the pair (lower index) is an oxygen (the reference's
 *                   test name(i) == "O" .and. name(i) == "O", :293).
 * Distances use the minimum image when periodic (box_image,
 * src/box_image.f90:36-64).  Literals the reference writes as default REAL
 * (k_rr, eps_OO) are widened from binary32 as the compiled reference sees
 * them.
 * Pinned by: central differences of E (plain Coulomb, where the gradient is
 * the derivative of the energy), and closed forms of single molecules and
 * pairs (tests/test_oracle_water.py).
 */
#include <math.h>
#include <stddef.h>
#include "oracle.h"

#define BOHR_A 0.52917721092
#define HARTREE_KCAL 627.5094743

void oracle_water_params(double par[11])
{
    par[0] = 1.0 / BOHR_A;                                   /* r0 OH */
    par[1] = 1.633 / BOHR_A;                                 /* r0 HH */
    par[2] = 101.9188 / HARTREE_KCAL;                        /* De */
    par[3] = 2.567 * BOHR_A;                                 /* a */
    par[4] = 328.645606 / HARTREE_KCAL * BOHR_A * BOHR_A;    /* k_HH */
    par[5] = -211.4672 / HARTREE_KCAL * BOHR_A * BOHR_A;     /* k_rtheta */
    par[6] = (double)111.70765f / HARTREE_KCAL * BOHR_A * BOHR_A;   /* k_rr (default REAL) */
    par[7] = 0.41;                                           /* q_H */
    par[8] = -0.82;                                          /* q_O */
    par[9] = 3.166 / BOHR_A;                                 /* sigma_OO */
    par[10] = (double)0.1554f / HARTREE_KCAL;                /* eps_OO (default REAL) */
}

static double image(double d, double box)
{
    const double h = 0.5 * box;
    while (fabs(d) > h) d -= copysign(box, d);
    return d;
}

double oracle_water_eval(int nwater, const double *x, double *g, int periodic, double box, int zahn,
                         double zahn_a, double zahn_par, double coul_cut)
{
    double par[11];
    oracle_water_params(par);
    const double r0 = par[0], r0hh = par[1], de = par[2], a = par[3], kth = par[4], krt = par[5], krr = par[6];
    const double qh = par[7], qo = par[8], sig = par[9], eps = par[10];
    const int n = 3 * nwater;
    double e = 0.0;
    for (int i = 0; i < 3 * n; ++i) g[i] = 0.0;
    for (int w = 0; w < nwater; ++w) {
        const int o = 3 * w, h1 = o + 1, h2 = o + 2;
        double v1[3], v2[3], vh[3], r1 = 0, r2 = 0, rh = 0;
        for (int c = 0; c < 3; ++c) {
            v1[c] = x[3 * o + c] - x[3 * h1 + c];
            v2[c] = x[3 * o + c] - x[3 * h2 + c];
            vh[c] = x[3 * h1 + c] - x[3 * h2 + c];
            if (periodic) {
                v1[c] = image(v1[c], box);
                v2[c] = image(v2[c], box);
                vh[c] = image(vh[c], box);
            }
            r1 += v1[c] * v1[c];
            r2 += v2[c] * v2[c];
            rh += vh[c] * vh[c];
        }
        r1 = sqrt(r1);
        r2 = sqrt(r2);
        rh = sqrt(rh);
        const double d1 = r1 - r0, d2 = r2 - r0, dh = rh - r0hh;
        const double e1 = exp(a * d1), e2 = exp(a * d2);
        e += de * (1.0 - e1) * (1.0 - e1) + de * (1.0 - e2) * (1.0 - e2);
        const double f1 = -2.0 * a * de * e1 * (1.0 - e1) / r1;    /* dE/dr1 / r1 */
        const double f2 = -2.0 * a * de * e2 * (1.0 - e2) / r2;
        e += 0.5 * kth * dh * dh;
        e += krt * dh * (d1 + d2);
        e += krr * d1 * d2;
        /* dE/dr of each distance, then along its vector */
        const double gr1 = f1 * r1 + krt * dh + krr * d2;           /* dE/dr1 */
        const double gr2 = f2 * r2 + krt * dh + krr * d1;
        const double grh = kth * dh + krt * (d1 + d2);              /* dE/drHH */
        for (int c = 0; c < 3; ++c) {
            const double a1 = gr1 * v1[c] / r1, a2 = gr2 * v2[c] / r2, ah = grh * vh[c] / rh;
            g[3 * o + c] += a1 + a2;
            g[3 * h1 + c] += -a1 + ah;
            g[3 * h2 + c] += -a2 - ah;
        }
    }
    for (int i = 0; i < n - 1; ++i)
        for (int j = i + 1; j < n; ++j) {
            if (i / 3 == j / 3) continue;
            double d[3], r2 = 0.0;
            for (int c = 0; c < 3; ++c) {
                d[c] = x[3 * i + c] - x[3 * j + c];
                if (periodic) d[c] = image(d[c], box);
                r2 += d[c] * d[c];
            }
            const double r = sqrt(r2);
            if (periodic && !(r < coul_cut)) continue;
            const double qi = (i % 3 == 0) ? qo : qh, qj = (j % 3 == 0) ? qo : qh;
            const double oner = 1.0 / r;
            const double e0 = zahn ? qi * qj * (erfc(zahn_a * r) * oner - zahn_par * (r - coul_cut)) : qi * qj * oner;
            e += e0;
            for (int c = 0; c < 3; ++c) {
                const double gv = e0 * oner * oner * d[c];
                g[3 * i + c] -= gv;
                g[3 * j + c] += gv;
            }
            if (i % 3 == 0) {
                const double s6 = pow(sig * oner, 6);
                e += 4.0 * eps * (s6 * s6 - s6);
                for (int c = 0; c < 3; ++c) {
                    const double gv = 24.0 * eps * oner * oner * d[c] * s6 * (2.0 * s6 - 1.0);
                    g[3 * i + c] -= gv;
                    g[3 * j + c] += gv;
                }
            }
        }
    return e;
}
