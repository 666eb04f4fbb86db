/*
 * oracle/pes_oh3.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * CPU restatement of the OH + H2 surface of Schatz and Elgersma (Chem. Phys.
 * Lett. 73, 21 (1980)) as the reference evaluates it in src/egrad_oh3.f
 * (POTLIB form).  Only tests/ may load this file's library; the product path
 * (caracal_amd) never calls it.
 *
 * Atoms: 1 = O, 2 = H1 (the hydroxyl hydrogen), 3 = H2, 4 = H3.  The six
 * distances (src/egrad_oh3.f:51-91): R1 = O-H1, R2 = O-H2, R3 = O-H3,
 * R4 = H1-H2, R5 = H1-H3, R6 = H2-H3.  V = three Morse pairs + a LEPS term
 * over (R2, R3, R6) + two H2O terms over (R1, R2, R4) and (R1, R3, R5) + a
 * four-body Gaussian over (R2, R3, R4, R5), shifted so that OH(re) + H2(re)
 * far apart is zero (pot_oh3, src/egrad_oh3.f:249-400).  Every constant
 * carries a D exponent (src/egrad_oh3.f:620-633), so no binary32 widening.
 *
 * Reference quirk kept: the three-body, H2O and four-body routines share one
 * derivative array (COMMON /POT2CM_oh3/).  When an H2O switching factor
 * Q(i) is exactly zero (x >= 43, an O-H distance above ~37.6 bohr), VH2O
 * leaves DEDR(i) unset (src/egrad_oh3.f:590-594), so the value of the
 * routine called before it stands in its place.
 *
 * Parity: pinned against the compiled reference (oracle/_ref, built from
 * src/egrad_oh3.f and src/util_oh3.f by oracle/Makefile.ref), by central
 * differences of V, by the H2 <-> H3 exchange symmetry and by the closed-form
 * asymptote V = 0 of OH(re) + H2(re) (tests/test_oracle_oh3.py).
 */
#include <math.h>
#include <string.h>
#include "oracle.h"

/* BLOCK DATA PTPACM_oh3 (src/egrad_oh3.f:620-633) */
static const double DE[3] = {0.148201, 0.0275690, 0.151548};
static const double BETA[3] = {1.260580, 0.924180, 1.068620};
static const double RE[3] = {1.863300, 2.907700, 1.428600};
static const double SATO = 0.10;
static const double GAM[3] = {2.399700, 1.058350, 2.399700};
static const double REOH = 1.808090, REHH = 2.861590;
static const double CON[7] = {-.0015920, 0.026963, 0.0014689, 0.080011, 0.085816, -0.063179, 0.101380};
static const double ALP[4] = {4.773, 7.14, 2.938, 5.28};
static const double CLAM[4] = {0.10, 0.10, 0.20, 0.03};
static const double ACON[2] = {0.10, 0.009};

static double vmor(double d, double b, double t, double r)
{
    const double u = 1.0 - exp(-b * (r - t));
    return d * (u * u);
}

static double dvmor(double d, double b, double t, double r)
{
    return 2.0 * b * d * (1.0 - exp(-b * (r - t))) * exp(-b * (r - t));
}

/* V3POT_oh3 (src/egrad_oh3.f:438-470): LEPS over r[0..2]; (DE, BETA, RE)
 * of pair 1 for the first two distances, of pair 3 for the third. */
static double v3pot(const double r[3], double dedr[3])
{
    const double de[3] = {DE[0], DE[0], DE[2]}, be[3] = {BETA[0], BETA[0], BETA[2]},
                 re[3] = {RE[0], RE[0], RE[2]};
    double x[3], exch[3], b[3], op3z[3], zp3[3];
    double energy = 0.0;
    for (int i = 0; i < 3; ++i) {
        const double z = SATO, zpo = 1.0 + z;
        op3z[i] = 1.0 + 3.0 * z;
        const double top3z = 2.0 * op3z[i];
        zp3[i] = z + 3.0;
        const double tzp3 = 2.0 * zp3[i];
        const double do4z = de[i] / 4.0 / zpo;
        b[i] = be[i] * do4z * 2.0;
        x[i] = exp(-be[i] * (r[i] - re[i]));
        const double coul = do4z * (zp3[i] * x[i] - top3z) * x[i];
        exch[i] = do4z * (op3z[i] * x[i] - tzp3) * x[i];
        energy = energy + coul;
    }
    const double d12 = exch[0] - exch[1], d23 = exch[1] - exch[2], d31 = exch[2] - exch[0];
    const double rad = sqrt(d12 * d12 + d23 * d23 + d31 * d31);
    energy = energy - rad / sqrt(2.0);
    const double s = exch[0] + exch[1] + exch[2];
    for (int i = 0; i < 3; ++i)
        dedr[i] = b[i] * x[i] * ((3.0 * exch[i] - s) / sqrt(2.0) * (op3z[i] * x[i] - zp3[i]) / rad
                                 - zp3[i] * x[i] + op3z[i]);
    return energy;
}

/* V4POT_oh3 (src/egrad_oh3.f:476-510) */
static double v4pot(const double r[4], double dedr[4])
{
    const double *a = ALP, *c = CLAM, *cof = ACON;
    const double e1 = r[0] - a[0], e2 = r[1] - a[0], e3 = r[2] - a[2], e4 = r[3] - a[2];
    const double f1 = r[0] - a[1], f2 = r[1] - a[1], f3 = r[2] - a[3], f4 = r[3] - a[3];
    const double t1 = exp(-c[0] * (e1 * e1) - c[0] * (e2 * e2) - c[2] * (e3 * e3) - c[2] * (e4 * e4)) * cof[0];
    const double t2 = exp(-c[1] * (f1 * f1) - c[1] * (f2 * f2) - c[3] * (f3 * f3) - c[3] * (f4 * f4)) * cof[1];
    dedr[0] = -2.0 * (t1 * c[0] * e1 + t2 * c[1] * f1);
    dedr[1] = -2.0 * (t1 * c[0] * e2 + t2 * c[1] * f2);
    dedr[2] = -2.0 * (t1 * c[2] * e3 + t2 * c[3] * f3);
    dedr[3] = -2.0 * (t1 * c[2] * e4 + t2 * c[3] * f4);
    return t1 + t2;
}

/* VH2O_oh3 (src/egrad_oh3.f:516-598) over (r_OH, r_OH', r_HH'); dedr holds
 * the shared array on entry (see the quirk above). */
static double vh2o(const double r[3], double dedr[3])
{
    const double xmax1 = 15.0, xmax2 = 43.0;
    double s[3], q[3], dq[3], x[3];
    s[0] = r[0] - REOH;
    s[2] = r[1] - REOH;
    s[1] = r[2] - REHH;
    for (int i = 0; i < 3; ++i) {
        x[i] = 0.5 * GAM[i] * s[i];
        q[i] = 1.0 - tanh(x[i]);
        if (x[i] >= xmax1) {
            if (x[i] >= xmax2) {
                q[i] = 0.0;
                dq[i] = 0.0;
                continue;
            }
            q[i] = 2.0 / (1.0 + exp(2.0 * x[i]));
        }
        const double ch = cosh(x[i]);
        dq[i] = -0.5 * GAM[i] / (ch * ch);
    }
    const double *c = CON;
    const double p = c[0] + c[1] * (s[0] + s[2]) + c[2] * s[1] + 0.5 * c[3] * (s[0] * s[0] + s[2] * s[2])
                     + 0.5 * c[4] * s[1] * s[1] + c[5] * s[1] * (s[0] + s[2]) + c[6] * s[0] * s[2];
    const double e = q[0] * q[1] * q[2] * p;
    double dp[3];
    dp[0] = c[1] + c[3] * s[0] + c[5] * s[1] + c[6] * s[2];
    dp[1] = c[2] + c[4] * s[1] + c[5] * (s[0] + s[2]);
    dp[2] = c[1] + c[3] * s[2] + c[5] * s[1] + c[6] * s[0];
    for (int i = 0; i < 3; ++i) {
        if (q[i] == 0.0) continue;
        const double trm1 = dq[i] / q[i];
        dedr[i] = e * (trm1 + (dp[i] / p));
    }
    const double t = dedr[1];
    dedr[1] = dedr[2];
    dedr[2] = t;
    return e;
}

void oracle_oh3_pot(const double r[6], double *vtot, double dvdr[6])
{
    double v = vmor(DE[0], BETA[0], RE[0], r[0]) + vmor(DE[1], BETA[1], RE[1], r[3])
               + vmor(DE[1], BETA[1], RE[1], r[4]);
    for (int i = 0; i < 6; ++i) dvdr[i] = 0.0;
    dvdr[0] = dvdr[0] + dvmor(DE[0], BETA[0], RE[0], r[0]);
    dvdr[3] = dvdr[3] + dvmor(DE[1], BETA[1], RE[1], r[3]);
    dvdr[4] = dvdr[4] + dvmor(DE[1], BETA[1], RE[1], r[4]);
    double shared[4];                      /* COMMON /POT2CM_oh3/ DEDR(1:3) */
    const double r3[3] = {r[1], r[2], r[5]};
    v = v + v3pot(r3, shared);
    dvdr[1] = dvdr[1] + shared[0];
    dvdr[2] = dvdr[2] + shared[1];
    dvdr[5] = dvdr[5] + shared[2];
    const double ra[3] = {r[0], r[1], r[3]};
    v = v + vh2o(ra, shared);
    dvdr[0] = dvdr[0] + shared[0];
    dvdr[1] = dvdr[1] + shared[1];
    dvdr[3] = dvdr[3] + shared[2];
    const double rb[3] = {r[0], r[2], r[4]};
    v = v + vh2o(rb, shared);
    dvdr[0] = dvdr[0] + shared[0];
    dvdr[2] = dvdr[2] + shared[1];
    dvdr[4] = dvdr[4] + shared[2];
    const double r4[4] = {r[1], r[2], r[3], r[4]};
    v = v + v4pot(r4, shared);
    dvdr[1] = dvdr[1] + shared[0];
    dvdr[2] = dvdr[2] + shared[1];
    dvdr[3] = dvdr[3] + shared[2];
    dvdr[4] = dvdr[4] + shared[3];
    *vtot = v - 2.0 * DE[1] + DE[2];
}

/* egrad_oh3 (src/egrad_oh3.f:33-190): distances, pot_oh3, and the chain rule
 * onto the Cartesian gradient, term by term in the reference's order.  Each
 * bead's gradient is its own (the reference clears the whole array inside
 * its bead loop and is called one bead at a time). */
void oracle_egrad_oh3(const double *q, int natoms, int nbeads, double *v, double *dvdq)
{
    static const int pa[6] = {1, 2, 3, 2, 3, 3}, pb[6] = {0, 0, 0, 1, 1, 2};
    (void)natoms;
    for (int k = 0; k < nbeads; ++k) {
        const double *x = q + 12 * k;
        double *g = dvdq + 12 * k;
        double d[6][3], r[6], dvdr[6];
        for (int m = 0; m < 6; ++m) {
            for (int c = 0; c < 3; ++c) d[m][c] = x[3 * pa[m] + c] - x[3 * pb[m] + c];
            r[m] = sqrt(d[m][0] * d[m][0] + d[m][1] * d[m][1] + d[m][2] * d[m][2]);
        }
        oracle_oh3_pot(r, &v[k], dvdr);
        memset(g, 0, 12 * sizeof(double));
        /* atom order of the reference's accumulation: O (R1, R2, R3), H1 (R1,
         * R4, R5), H2 (R2, R4, R6), H3 (R3, R5, R6) */
        static const int terms[4][3] = {{0, 1, 2}, {0, 3, 4}, {1, 3, 5}, {2, 4, 5}};
        for (int at = 0; at < 4; ++at)
            for (int t = 0; t < 3; ++t) {
                const int m = terms[at][t];
                const double sgn = pa[m] == at ? 1.0 : -1.0;
                for (int c = 0; c < 3; ++c) {
                    const double u = dvdr[m] * d[m][c] / r[m];
                    g[3 * at + c] = sgn > 0 ? g[3 * at + c] + u : g[3 * at + c] - u;
                }
            }
    }
}
