/*
 * oracle/pes_brh2.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * CPU restatement of the BrH2 DIM-3C surface (diatomics-in-molecules plus a
 * three-centre term; D. C. Clary, Chem. Phys. 71, 117 (1982); I. Last and
 * M. Baer) as the reference evaluates it in src/egrad_brh2.f (POTLIB form).
 * Only tests/ may load this file's library; the product path (caracal_amd)
 * never calls it.
 *
 * Atoms (src/egrad_brh2.f:46-64): 1 = Ha, 2 = Br, 3 = Hc; the surface's
 * distances are R1 = Br-Ha, R2 = Br-Hc, R3 = Ha-Hc (POT_brh2 reads them as
 * R(1), R(3), R(2), :217-219).  V = the lowest eigenvalue of a 4 x 4 DIM
 * matrix (the H2 singlet/triplet curves and the HBr 1S, 3S, 1P, 3P curves
 * mixed by the two Br-H angles, :217-381) + a three-centre overlap term + DHH,
 * so that Br far from H2(re) is zero (:413-432).  Derivatives of the
 * eigenvalue follow Hellmann-Feynman, u^T (dH/dR) u with u its eigenvector
 * (:386-410).
 *
 * Eigen-solve: the reference calls EISPACK (RSP: TRED3, TQL2, TRBAK3,
 * :485-1059, called at :386) for all four eigenpairs and takes the lowest.
 * The oracle uses the cyclic Jacobi method instead (plane rotations that zero each
 * off-diagonal element in turn, t = sgn(theta) / (|theta| + sqrt(theta^2 + 1)),
 * theta = (a_qq - a_pp) / (2 a_pq); sweeps until the off-diagonal elements
 * sum to at most 2^-60 of the diagonal's magnitudes): the lowest eigenvalue and the quadratic form u^T D u
 * do not depend on the method beyond rounding (nor on the sign of u).
 *
 * Parity: pinned against the compiled reference (oracle/_ref, built from
 * src/egrad_brh2.f by oracle/Makefile.ref), by central differences of V, by
 * the Ha <-> Hc exchange symmetry and by the closed forms of the separated
 * fragments Br + H2 (V = V1HH(r) + DHH) and H + HBr (V = V1S(r) + DHH)
 * (tests/test_oracle_brh2.py).
 */
#include <math.h>
#include "oracle.h"

/* BLOCK DATA PTPARM (src/egrad_brh2.f:466-483), all with a D exponent */
static const double EPS = 1e-6, RHH = 1.4016, AHH = 1.0291, DHH = 0.17447, BHH = 0.018;
static const double XIH = 1.0, XIX = 1.6, G0 = 0.22, ALFW = 1.0;
static const double RHX = 2.673, AHX = 0.957, DHX = 0.1439, BHX = 0.012;
static const double ETAHH = 0.393764, ETA3S = 0.322, ETA1P = 0.20286, ETA3P = 0.1771;

/* HBr curves (VHX, src/egrad_brh2.f:442-463): v = (1S, 3S, 1P, 3P) and r-derivatives */
static void vhx(double r, double v[4], double dv[4])
{
    const double rdif = r - RHX, rdif2 = rdif * rdif;
    const double ex1 = exp(-AHX * rdif), ex2 = exp(-BHX * rdif2 * rdif);
    double t1 = DHX * ex1 * ex2;
    const double v1 = t1 * (ex1 + 2.0);
    v[0] = t1 * (ex1 - 2.0);
    v[1] = ETA3S * v1;
    v[2] = ETA1P * v1;
    v[3] = ETA3P * v1;
    t1 = 2.0 * AHX * t1;
    const double t2 = 3.0 * BHX * rdif2;
    const double dv1 = -t1 * (1.0 + ex1) - t2 * v1;
    dv[0] = t1 * (1.0 - ex1) - t2 * v[0];
    dv[1] = ETA3S * dv1;
    dv[2] = ETA1P * dv1;
    dv[3] = ETA3P * dv1;
}

/* packed lower triangle, row-wise: (i, j), i >= j, 0-based -> i (i + 1) / 2 + j */
static int pk(int i, int j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

/* Lowest eigenpair of the symmetric 4 x 4 matrix h (packed) by cyclic Jacobi. */
static double lowest_eigenpair(const double h[10], double u[4])
{
    double a[4][4], v[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            a[i][j] = h[pk(i, j)];
            v[i][j] = i == j ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int p = 0; p < 4; ++p) {
            diag += fabs(a[p][p]);
            for (int q = p + 1; q < 4; ++q) off += fabs(a[p][q]);
        }
        /* converged: what is left moves an eigenvalue by O(off^2 / gap) and
           the eigenvectors by O(off / gap), far below rounding */
        if (off <= 0x1p-60 * diag) break;
        for (int p = 0; p < 4; ++p)
            for (int q = p + 1; q < 4; ++q) {
                const double apq = a[p][q];
                if (apq == 0.0) continue;
                const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
                double t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
                if (theta < 0.0) t = -t;
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                a[p][p] -= t * apq;
                a[q][q] += t * apq;
                a[p][q] = a[q][p] = 0.0;
                for (int r = 0; r < 4; ++r) {
                    if (r != p && r != q) {
                        const double arp = a[r][p], arq = a[r][q];
                        a[r][p] = a[p][r] = c * arp - s * arq;
                        a[r][q] = a[q][r] = s * arp + c * arq;
                    }
                    const double vrp = v[r][p], vrq = v[r][q];
                    v[r][p] = c * vrp - s * vrq;
                    v[r][q] = s * vrp + c * vrq;
                }
            }
    }
    int m = 0;
    for (int i = 1; i < 4; ++i)
        if (a[i][i] < a[m][m]) m = i;
    for (int i = 0; i < 4; ++i) u[i] = v[i][m];
    return a[m][m];
}

/* POT_brh2 (src/egrad_brh2.f:207-441): r[0] = R(1) = Br-Ha, r[1] = R(2) = Ha-Hc,
 * r[2] = R(3) = Br-Hc; dedr in the same order. */
void oracle_brh2_pot(const double r[3], double *energy, double dedr[3])
{
    /* initialize_brh2's derived constants (src/egrad_brh2.f:174-180) */
    const double xib = 0.5 * (XIH + XIX), tt = xib * RHX, c3 = 1.0 / 3.0;
    const double denom = 2.0 * RHX * (1.0 + tt * (1.0 + c3 * tt)) * exp(-tt);
    const double g = G0 / RHX, rt34 = 0.25 * sqrt(3.0), rt38 = 0.5 * rt34;

    double r1 = r[0], r2 = r[2], r3 = r[1];
    if (r1 > r2 + r3) r1 = r2 + r3;
    if (r2 > r1 + r3) r2 = r1 + r3;
    if (r3 > r1 + r2) r3 = r1 + r2;
    const double r1s = r1 * r1, r2s = r2 * r2, r3s = r3 * r3;
    const double r12 = r1 * r2, r13 = r1 * r3, r23 = r2 * r3;
    const double c3x = (r1s + r2s - r3s) * 0.5, c2x = (r1s + r3s - r2s) * 0.5, c1x = (r2s + r3s - r1s) * 0.5;

    /* angle at Ha (between Br-Ha and Ha-Hc) */
    double csa = c2x / r13;
    if (fabs(csa) > 1.0) csa = csa > 0.0 ? 1.0 : -1.0;
    const double csa2 = csa * csa, sna2 = 1.0 - csa2;
    const int lcol = sna2 < EPS;
    const double sna = sqrt(sna2);
    double t22 = 2.0 * csa, t11 = lcol ? 0.0 : 2.0 * (sna2 - csa2) / sna;
    const double sn2a = t22 * sna;
    double t = c3x / (r1 * r13);
    const double dcsa21 = t22 * t, dsn2a1 = t11 * t;
    t = -r2 / r13;
    const double dcsa22 = t22 * t, dsn2a2 = t11 * t;
    t = c1x / (r13 * r3);
    const double dcsa23 = t22 * t, dsn2a3 = t11 * t;

    /* angle at Hc */
    double csb = c1x / r23;
    if (fabs(csb) > 1.0) csb = csb > 0.0 ? 1.0 : -1.0;
    const double csb2 = csb * csb, snb2 = 1.0 - csb2, snb = sqrt(snb2);
    t11 = lcol ? 0.0 : 2.0 * (snb2 - csb2) / snb;
    t22 = 2.0 * csb;
    const double sn2b = t22 * snb;
    t = -r1 / r23;
    const double dcsb21 = t22 * t, dsn2b1 = t11 * t;
    t = c3x / (r2 * r23);
    const double dcsb22 = t22 * t, dsn2b2 = t11 * t;
    t = c2x / (r23 * r3);
    const double dcsb23 = t22 * t, dsn2b3 = t11 * t;

    /* cos^2 of the angle at Br and its derivatives */
    t = c3x / r12;
    const double csg2 = t * t;
    t = 2.0 * t;
    const double dcsg21 = t * c2x / (r1 * r12), dcsg22 = t * c1x / (r12 * r2), dcsg23 = -t * r3 / r12;

    /* H2 singlet and triplet */
    double rdif = r3 - RHH;
    double ex1 = exp(-AHH * rdif);
    const double rdif2 = rdif * rdif, ex2 = exp(-BHH * rdif * rdif2);
    double a1 = DHH * ex1 * ex2;
    const double v1hh = a1 * (ex1 - 2.0), v3hh = ETAHH * a1 * (ex1 + 2.0);
    a1 = 2.0 * AHH * a1;
    const double a2 = 3.0 * BHH * rdif2;
    const double dv1hh = a1 * (1.0 - ex1) - a2 * v1hh;
    const double dv3hh = -a1 * ETAHH * (1.0 + ex1) - a2 * v3hh;

    double w1[4], dw1[4], w2[4], dw2[4];
    vhx(r1, w1, dw1);
    vhx(r2, w2, dw2);
    /* singlet/triplet combinations of pair n (1 = Br-Ha, 2 = Br-Hc) */
    const double s11 = w1[0] + 3.0 * w1[1], s12 = w2[0] + 3.0 * w2[1];
    const double s21 = 3.0 * w1[0] + w1[1], s22 = 3.0 * w2[0] + w2[1];
    const double s31 = w1[0] - w1[1], s32 = w2[0] - w2[1];
    const double p11 = w1[2] + 3.0 * w1[3], p12 = w2[2] + 3.0 * w2[3];
    const double p21 = 3.0 * w1[2] + w1[3], p22 = 3.0 * w2[2] + w2[3];
    const double p31 = w1[2] - w1[3], p32 = w2[2] - w2[3];
    const double ds11 = dw1[0] + 3.0 * dw1[1], ds12 = dw2[0] + 3.0 * dw2[1];
    const double ds21 = 3.0 * dw1[0] + dw1[1], ds22 = 3.0 * dw2[0] + dw2[1];
    const double ds31 = dw1[0] - dw1[1], ds32 = dw2[0] - dw2[1];
    const double dp11 = dw1[2] + 3.0 * dw1[3], dp12 = dw2[2] + 3.0 * dw2[3];
    const double dp21 = 3.0 * dw1[2] + dw1[3], dp22 = 3.0 * dw2[2] + dw2[3];
    const double dp31 = dw1[2] - dw1[3], dp32 = dw2[2] - dw2[3];

    /* DIM matrix, packed (src/egrad_brh2.f:322-341) */
    double h[10], dh1[10], dh2[10], dh3[10];
    h[0] = v1hh + 0.25 * (s11 * csa2 + p11 * sna2 + s12 * csb2 + p12 * snb2);
    h[2] = v1hh + 0.25 * (s11 * sna2 + p11 * csa2 + s12 * snb2 + p12 * csb2);
    h[5] = v3hh + 0.25 * (s21 * csa2 + p21 * sna2 + s22 * csb2 + p22 * snb2);
    h[9] = v3hh + 0.25 * (s21 * sna2 + p21 * csa2 + s22 * snb2 + p22 * csb2);
    const double u11 = s11 - p11, u12 = s12 - p12;
    h[1] = lcol ? 0.0 : 0.125 * (u11 * sn2a - u12 * sn2b);
    h[3] = rt34 * (s31 * csa2 + p31 * sna2 - s32 * csb2 - p32 * snb2);
    const double u31 = s31 - p31, u32 = s32 - p32;
    h[4] = lcol ? 0.0 : rt38 * (u31 * sn2a + u32 * sn2b);
    h[6] = h[4];
    h[7] = rt34 * (s31 * sna2 + p31 * csa2 - s32 * snb2 - p32 * csb2);
    const double u21 = s21 - p21, u22 = s22 - p22;
    h[8] = lcol ? 0.0 : 0.125 * (u21 * sn2a - u22 * sn2b);

    /* its derivatives by R1, R2, R3 (:343-381) */
    t = u11 * dcsa21 + u12 * dcsb21;
    dh1[0] = 0.25 * (ds11 * csa2 + dp11 * sna2 + t);
    dh1[1] = 0.125 * ((ds11 - dp11) * sn2a + u11 * dsn2a1 - u12 * dsn2b1);
    dh1[2] = 0.25 * (ds11 * sna2 + dp11 * csa2 - t);
    t = u11 * dcsa22 + u12 * dcsb22;
    dh2[0] = 0.25 * (ds12 * csb2 + dp12 * snb2 + t);
    dh2[1] = 0.125 * (-(ds12 - dp12) * sn2b + u11 * dsn2a2 - u12 * dsn2b2);
    dh2[2] = 0.25 * (ds12 * snb2 + dp12 * csb2 - t);
    t = 0.25 * (u11 * dcsa23 + u12 * dcsb23);
    dh3[0] = dv1hh + t;
    dh3[1] = 0.125 * (u11 * dsn2a3 - u12 * dsn2b3);
    dh3[2] = dv1hh - t;
    t = u21 * dcsa21 + u22 * dcsb21;
    dh1[5] = 0.25 * (ds21 * csa2 + dp21 * sna2 + t);
    dh1[8] = 0.125 * ((ds21 - dp21) * sn2a + u21 * dsn2a1 - u22 * dsn2b1);
    dh1[9] = 0.25 * (ds21 * sna2 + dp21 * csa2 - t);
    t = u21 * dcsa22 + u22 * dcsb22;
    dh2[5] = 0.25 * (ds22 * csb2 + dp22 * snb2 + t);
    dh2[8] = 0.125 * ((dp22 - ds22) * sn2b + u21 * dsn2a2 - u22 * dsn2b2);
    dh2[9] = 0.25 * (ds22 * snb2 + dp22 * csb2 - t);
    t = 0.25 * (u21 * dcsa23 + u22 * dcsb23);
    dh3[5] = dv3hh + t;
    dh3[8] = 0.125 * (u21 * dsn2a3 - u22 * dsn2b3);
    dh3[9] = dv3hh - t;
    t = u31 * dcsa21 - u32 * dcsb21;
    dh1[3] = rt34 * (ds31 * csa2 + dp31 * sna2 + t);
    dh1[4] = rt38 * ((ds31 - dp31) * sn2a + u31 * dsn2a1 + u32 * dsn2b1);
    dh1[6] = dh1[4];
    dh1[7] = rt34 * (ds31 * sna2 + dp31 * csa2 - t);
    t = u31 * dcsa22 - u32 * dcsb22;
    dh2[3] = rt34 * (-ds32 * csb2 - dp32 * snb2 + t);
    dh2[4] = rt38 * ((ds32 - dp32) * sn2b + u31 * dsn2a2 + u32 * dsn2b2);
    dh2[6] = dh2[4];
    dh2[7] = rt34 * (-ds32 * snb2 - dp32 * csb2 - t);
    t = u31 * dcsa23 - u32 * dcsb23;
    dh3[3] = rt34 * t;
    dh3[4] = rt38 * (u31 * dsn2a3 + u32 * dsn2b3);
    dh3[6] = dh3[4];
    dh3[7] = -dh3[3];

    double u[4];
    const double e1 = lowest_eigenpair(h, u);
    double d1 = 0.0, d2 = 0.0, d3 = 0.0;
    for (int l = 0; l < 4; ++l)
        for (int k = 0; k < 4; ++k) {
            const double w = u[l] * u[k];
            d1 += w * dh1[pk(l, k)];
            d2 += w * dh2[pk(l, k)];
            d3 += w * dh3[pk(l, k)];
        }

    /* three-centre term (:413-432) */
    t = XIH * r3;
    ex1 = exp(-t);
    const double shh = (1.0 + t * (1.0 + c3 * t)) * ex1;
    const double dshh = -XIH * t * (1.0 + t) * c3 * ex1;
    t = xib * r1;
    ex1 = exp(-t);
    const double shx1 = r1 * (1.0 + t * (1.0 + c3 * t)) * ex1 / denom;
    const double dshx1 = (1.0 + t * (1.0 - c3 * t * t)) * ex1 / denom;
    t = xib * r2;
    ex1 = exp(-t);
    const double shx2 = r2 * (1.0 + t * (1.0 + c3 * t)) * ex1 / denom;
    const double dshx2 = (1.0 + t * (1.0 - c3 * t * t)) * ex1 / denom;
    rdif = r1 - r2;
    const double b1 = g * exp(-ALFW * rdif * rdif);
    const double b2 = shh * (shx1 + shx2) + shx1 * shx2;
    *energy = e1 + b2 * b1 * csg2 + DHH;
    const double b3 = 2.0 * ALFW * rdif * csg2;
    dedr[0] = d1 + (dshx1 * (shh + shx2) * csg2 + b2 * (dcsg21 - b3)) * b1;
    dedr[2] = d2 + (dshx2 * (shh + shx1) * csg2 + b2 * (dcsg22 + b3)) * b1;
    dedr[1] = d3 + (dshh * (shx1 + shx2) * csg2 + b2 * dcsg23) * b1;
}

/* egrad_brh2 (src/egrad_brh2.f:33-79): q[bead][atom][xyz], atoms Ha, Br, Hc */
void oracle_egrad_brh2(const double *q, int natoms, int nbeads, double *v, double *dvdq)
{
    for (int k = 0; k < nbeads; ++k) {
        const double *x = q + (size_t)k * natoms * 3;
        double *gk = dvdq + (size_t)k * natoms * 3;
        double ab[3], ac[3], bc[3];
        for (int c = 0; c < 3; ++c) {
            ab[c] = x[3 + c] - x[c];
            ac[c] = x[c] - x[6 + c];
            bc[c] = x[6 + c] - x[3 + c];
        }
        const double r[3] = {sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]),
                             sqrt(ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2]),
                             sqrt(bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2])};
        double d[3];
        oracle_brh2_pot(r, &v[k], d);
        for (int c = 0; c < 3; ++c) {
            gk[c] = d[1] * ac[c] / r[1] - d[0] * ab[c] / r[0];
            gk[3 + c] = d[0] * ab[c] / r[0] - d[2] * bc[c] / r[2];
            gk[6 + c] = d[2] * bc[c] / r[2] - d[1] * ac[c] / r[1];
        }
    }
}
