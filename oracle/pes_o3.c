/*
 * oracle/pes_o3.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * CPU restatement of the O + O2 surface of Z. Varga, Y. Paukku and D. G.
 * Truhlar (J. Chem. Phys. 147, 154312 (2017); the 1 1A'' singlet) as the
 * reference evaluates it in src/egrad_o3.f (POTLIB form).  Only tests/ may
 * load this file's library; the product path (caracal_amd) never calls it.
 *
 * Units inside the surface: angstrom and kcal/mol (pot_o3, src/egrad_o3.f:
 * 83-132: x_A = 0.52917721092 x_bohr; E = v 0.159360144e-2 - 0.19172848
 * hartree; dE/dx = dv/dx_A 0.843297564e-3 hartree/bohr, a rounded factor
 * that is Econv * Cconv only to 1e-9, so the gradient is the energy's
 * derivative only to that relative precision).  Distances (:187-211):
 * R1 = O1-O2, R2 = O1-O3, R3 = O2-O3.  v = 2 De (240.486) + sum of the
 * diatomic curve over the three R (ev2gm2_o3, :503-565: eight Gaussians in
 * r^2 times 0.627509523475149) + a D3(BJ) dispersion over the three pairs
 * (d3disp_o3 / edisp_o3, :906-1030) + sum_k C_k B_k, B the 56 permutationally
 * invariant polynomials of the mixed exponential-Gaussian variables
 * m_i = exp(-(R_i - ra)/a - (R_i - rb)^2/ab) that are neither constant nor
 * two-body (EvMorse_o3 .. evbas_o3, :323-501).
 *
 * Derivatives: the reference differentiates the polynomial recursion line by
 * line (EvdPdR_o3, :759-851); the oracle carries dual numbers (value and the
 * three partials by R1, R2, R3) through the same recursion.
 *
 * Reference quirks kept: d3disp_o3 converts its distance argument to bohr
 * and back in place, and it is called on the common-block distances, so the
 * gradient sees R after one (its two-body term) and two (the MEG
 * derivatives and dR/dx) such round trips (:227-241, :581-590); the D3
 * r2r4 table is written without a kind (default REAL), so oxygen's entry is
 * the binary32 value of 2.59361680 widened to binary64.
 *
 * Parity: pinned against the compiled reference (oracle/_ref, built from
 * src/egrad_o3.f by oracle/Makefile.ref), by central differences of V (to
 * the factor above), by the permutation symmetry of the three atoms and by
 * the separated limit: with O far from O2 every kept polynomial vanishes
 * (each has a factor from at least two of the three distances), so
 * E = (240.486 + V2(r) + the three pairs' dispersion) Econv + Eref
 * (tests/test_oracle_o3.py).
 */
#include <math.h>
#include "oracle.h"

/* BLOCK DATA prmt_o3 (src/egrad_o3.f:1057-1176) */
static const double MEG_A = 0.83, MEG_AB = 3.70, MEG_RA = 1.25, MEG_RB = 1.13;
static const double COEF[56] = {
    -0.128814549305e+03, 0.104229418850e+04, 0.811983220935e+03, -0.443324528752e+03,
    0.904506805268e+03, -0.501026918125e+04, 0.197209669844e+05, -0.251424247013e+05,
    -0.138013677810e+03, 0.169329202490e+05, 0.352627837493e+05, -0.334337897178e+05,
    0.720500009412e+05, 0.116986232065e+05, -0.521801943104e+04, -0.332486978745e+05,
    -0.177870892015e+05, 0.335720198273e+05, 0.268323174511e+05, -0.933618945467e+05,
    -0.592242307973e+04, 0.287777488764e+04, 0.393607079595e+05, -0.330171644074e+04,
    0.200362806379e+05, -0.975981166385e+04, -0.266133829509e+05, 0.746650532707e+05,
    0.120290055844e+05, -0.464904653691e+04, 0.244129022324e+04, -0.273870502550e+05,
    -0.122995471301e+05, 0.722408057250e+04, 0.290562738593e+05, -0.212778140565e+05,
    -0.206522536997e+05, 0.263237776823e+05, -0.370869661253e+05, -0.230814810543e+04,
    0.210942828415e+04, -0.177122132932e+04, 0.102466183681e+05, -0.556070962327e+03,
    0.150041418580e+05, -0.117657568996e+05, -0.442916346835e+04, 0.140353851796e+05,
    0.753137518090e+04, -0.911889476033e+04, 0.816658803687e+04, 0.316234339832e+04,
    -0.202814305330e+04, 0.791358291948e+03, 0.115450049995e+03, -0.158182849802e+04,
};

/* The polynomial recursion of EvPoly_o3 (src/egrad_o3.f:366-447) as a table:
 * p[k] = p[a] p[b] - p[s0] - p[s1] - p[s2] (in that order; -1 ends the list).
 * p[0] = 1, p[1] = m1 + m2 + m3, p[2] = m1 m2 + m1 m3 + m2 m3 and
 * p[4] = m1 m2 m3 come from the monomials. */
static const signed char REC[67][5] = {
    {-1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}, {1, 1, 2, 2, -1},
    {-1, -1, -1, -1, -1}, {1, 2, 4, 4, 4}, {1, 3, 5, -1, -1}, {1, 4, -1, -1, -1},
    {2, 2, 7, 7, -1}, {2, 3, 7, -1, -1}, {1, 6, 9, -1, -1}, {2, 4, -1, -1, -1},
    {3, 4, -1, -1, -1}, {1, 8, 11, -1, -1}, {2, 6, 12, -1, -1}, {1, 10, 14, -1, -1},
    {4, 4, -1, -1, -1}, {4, 5, -1, -1, -1}, {4, 6, -1, -1, -1}, {2, 8, 17, -1, -1},
    {1, 13, 17, 19, 19}, {2, 10, 18, -1, -1}, {1, 15, 21, -1, -1}, {1, 16, -1, -1, -1},
    {4, 8, -1, -1, -1}, {3, 11, 23, -1, -1}, {4, 10, -1, -1, -1}, {1, 19, 24, -1, -1},
    {6, 8, 23, -1, -1}, {2, 15, 26, -1, -1}, {1, 22, 29, -1, -1}, {2, 16, -1, -1, -1},
    {3, 16, -1, -1, -1}, {1, 24, 31, -1, -1}, {4, 14, -1, -1, -1}, {4, 15, -1, -1, -1},
    {2, 19, 33, -1, -1}, {3, 19, 31, -1, -1}, {8, 10, 32, -1, -1}, {2, 22, 35, -1, -1},
    {1, 30, 39, -1, -1}, {4, 16, -1, -1, -1}, {4, 17, -1, -1, -1}, {4, 19, -1, -1, -1},
    {6, 16, -1, -1, -1}, {4, 20, -1, -1, -1}, {4, 21, -1, -1, -1}, {4, 22, -1, -1, -1},
    {1, 36, 43, -1, -1}, {1, 37, 45, 48, -1}, {8, 15, 44, -1, -1}, {2, 30, 47, -1, -1},
    {1, 40, 51, -1, -1}, {1, 41, -1, -1, -1}, {4, 24, -1, -1, -1}, {3, 31, 53, -1, -1},
    {1, 43, 54, -1, -1}, {10, 16, -1, -1, -1}, {4, 28, -1, -1, -1}, {4, 29, -1, -1, -1},
    {4, 30, -1, -1, -1}, {2, 36, 56, -1, -1}, {3, 36, 54, -1, -1}, {10, 19, 53, -1, -1},
    {8, 22, 57, -1, -1}, {2, 40, 60, -1, -1}, {1, 52, 65, -1, -1},
};

/* evbas_o3 (:448-501): B keeps p[k] for every k but these eleven, in order */
static int dropped(int k)
{
    return k == 0 || k == 1 || k == 3 || k == 6 || k == 10 || k == 15 || k == 22 || k == 30 || k == 40 ||
           k == 52 || k == 66;
}

typedef struct { double v, d[3]; } dual3;

static dual3 dmul(dual3 a, dual3 b)
{
    dual3 r;
    r.v = a.v * b.v;
    for (int i = 0; i < 3; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}

static dual3 dsub(dual3 a, dual3 b)
{
    dual3 r;
    r.v = a.v - b.v;
    for (int i = 0; i < 3; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}

static dual3 dadd(dual3 a, dual3 b)
{
    dual3 r;
    r.v = a.v + b.v;
    for (int i = 0; i < 3; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}

/* the diatomic curve (ev2gm2_o3, imol 4) in kcal/mol and its r-derivative, r in angstrom */
static void v2(double r, double *v, double *dv)
{
    static const double alpha = 9.439784362354936e-1, beta = 1.262242998506810e0;
    static const double a[8] = {-1.488979427684798e3, 1.881435846488955e4, -1.053475425838226e5,
                                2.755135591229064e5,  -4.277588997761775e5, 4.404104009614092e5,
                                -2.946204062950765e5, 1.176861219078620e5};
    double s = 0.0, g = 0.0, bk = 1.0;
    for (int k = 0; k < 8; ++k) {
        const double e = exp(-alpha * bk * r * r);
        s += a[k] * e;
        g -= 2.0 * a[k] * alpha * bk * r * e;
        bk *= beta;
    }
    *v = s * 627.509523475149e-3;
    *dv = g * 627.509523475149e-3;
}

static const double AUTOANG = 0.52917726, AUTOKCAL = 627.509541;

/* D3(BJ) dispersion of one O-O pair at r (angstrom): energy in kcal/mol and its r-derivative */
static void disp(double r, double *e, double *de)
{
    const double s6 = 1.0, s8 = 2.0, a1 = 0.5299, a2 = 2.20, c6 = 12.8;
    const double r2r4 = (double)2.59361680f;
    const double c8 = 3.0 * c6 * r2r4 * r2r4;
    const double rb = r / AUTOANG;
    const double r0 = a1 * sqrt(c8 / c6) + a2;
    const double r02 = r0 * r0, r04 = r02 * r02, r06 = r04 * r02, r08 = r04 * r04;
    const double x2 = rb * rb, x4 = x2 * x2, x6 = x4 * x2, x8 = x4 * x4;
    const double d6 = x6 + r06, d8 = x8 + r08;
    const double e6 = c6 / d6, e8 = c8 / d8;
    const double e6dr = c6 * (-6.0 * x4 * rb) / (d6 * d6), e8dr = c8 * (-8.0 * x6 * rb) / (d8 * d8);
    *e = (-s6 * e6 - s8 * e8) * AUTOKCAL;
    *de = (-s6 * e6dr - s8 * e8dr) * AUTOKCAL / AUTOANG;
}

/* o3pes (:144-185): x in angstrom, v in kcal/mol, dvdx in kcal/(mol angstrom) */
static void o3pes(const double x[9], double *v, double dvdx[9])
{
    double r[3];
    const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};   /* R1 = |x1 - x0|, R2 = |x2 - x0|, R3 = |x1 - x2| */
    for (int i = 0; i < 3; ++i) {
        double s = 0.0;
        for (int c = 0; c < 3; ++c) {
            const double d = x[3 * pb[i] + c] - x[3 * pa[i] + c];
            s += d * d;
        }
        r[i] = sqrt(s);
    }
    /* energy with the distances as computed (EvV_o3, :213-273) */
    dual3 m[3];
    for (int i = 0; i < 3; ++i) {
        const double u = r[i] - MEG_RB;
        m[i].v = exp(-(r[i] - MEG_RA) / MEG_A - u * u / MEG_AB);
        m[i].d[0] = m[i].d[1] = m[i].d[2] = 0.0;
    }
    double e = 240.486;
    for (int i = 0; i < 3; ++i) {
        double vv, dv;
        v2(r[i], &vv, &dv);
        e += vv;
    }
    double ed = 0.0;
    for (int i = 0; i < 3; ++i) {
        double x6, d6;
        disp(r[i], &x6, &d6);
        ed += x6;
    }
    e += ed;
    /* d3disp's round trip through bohr, once for the energy, once for the gradient */
    double r1[3], r2[3];
    for (int i = 0; i < 3; ++i) {
        r1[i] = (r[i] / AUTOANG) * AUTOANG;
        r2[i] = (r1[i] / AUTOANG) * AUTOANG;
    }
    /* MEG derivatives at r2 (EvdMsdR_o3), the polynomial values at r */
    for (int i = 0; i < 3; ++i) {
        const double u = r2[i] - MEG_RB;
        m[i].d[i] = (-2.0 * u / MEG_AB - 1.0 / MEG_A) * exp(-(r2[i] - MEG_RA) / MEG_A - u * u / MEG_AB);
    }
    /* monomials (EvMono_o3): rm1 = m(R3), rm2 = m(R2), rm3 = m(R1) */
    const dual3 rm1 = m[2], rm2 = m[1], rm3 = m[0];
    const dual3 rm4 = dmul(rm1, rm2), rm5 = dmul(rm1, rm3), rm6 = dmul(rm2, rm3), rm7 = dmul(rm1, rm6);
    dual3 p[67];
    p[0].v = 1.0;
    p[0].d[0] = p[0].d[1] = p[0].d[2] = 0.0;
    p[1] = dadd(dadd(rm1, rm2), rm3);
    p[2] = dadd(dadd(rm4, rm5), rm6);
    p[4] = rm7;
    for (int k = 3; k < 67; ++k) {
        if (k == 4) continue;
        dual3 t = dmul(p[REC[k][0]], p[REC[k][1]]);
        for (int j = 2; j < 5 && REC[k][j] >= 0; ++j) t = dsub(t, p[REC[k][j]]);
        p[k] = t;
    }
    double dvdr[3];
    for (int i = 0; i < 3; ++i) {
        double vv, dv, x6, d6;
        v2(r1[i], &vv, &dv);
        disp(r1[i], &x6, &d6);
        dvdr[i] = dv + d6;
    }
    int n = 0;
    for (int k = 0; k < 67; ++k) {
        if (dropped(k)) continue;
        e += COEF[n] * p[k].v;
        for (int i = 0; i < 3; ++i) dvdr[i] += COEF[n] * p[k].d[i];
        ++n;
    }
    *v = e;
    /* chain rule with dR/dx at r2 (EvdRdX_o3, :634-690) */
    for (int i = 0; i < 9; ++i) dvdx[i] = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int c = 0; c < 3; ++c) {
            const double d = (x[3 * pa[i] + c] - x[3 * pb[i] + c]) / r2[i];
            dvdx[3 * pa[i] + c] += dvdr[i] * d;
            dvdx[3 * pb[i] + c] -= dvdr[i] * d;
        }
}

/* egrad_o3 (src/egrad_o3.f:33-56) with pot_o3 (:83-132): q[bead][atom][xyz] in bohr */
void oracle_egrad_o3(const double *q, int natoms, int nbeads, double *v, double *dvdq)
{
    const double cconv = 0.52917721092, econv = 0.159360144e-2, gconv = 0.843297564e-3, eref = -0.19172848;
    for (int k = 0; k < nbeads; ++k) {
        const double *xk = q + (size_t)k * natoms * 3;
        double x[9], g[9], e;
        for (int i = 0; i < 9; ++i) x[i] = xk[i] * cconv;
        o3pes(x, &e, g);
        v[k] = e * econv + eref;
        for (int i = 0; i < 9; ++i) dvdq[(size_t)k * natoms * 3 + i] = g[i] * gconv;
    }
}
