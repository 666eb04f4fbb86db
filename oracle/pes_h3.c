/*
 * oracle/pes_h3.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * CPU restatement of the BKMP2 H3 potential-energy surface as the reference
 * evaluates it in src/egrad_h3.f.  Only tests/, the smoke check and the CPU
 * baseline leg of bench.py may load this file's library; the product path
 * (caracal_amd) never calls it.
 *
 * Numerics follow the reference build exactly where it matters for parity:
 *  - every PARAMETER literal written without a D exponent is a default REAL,
 *    i.e. rounded to binary32 and then widened; SP() reproduces that;
 *  - sums are evaluated left to right in the order the reference writes them;
 *  - the unused second derivative of the H2 singlet curve is not evaluated
 *    (its result never reaches V or dV/dR in the reference).
 *
 * Parity: pinned against the compiled reference (oracle/_ref, built from
 * src/egrad_h3.f by oracle/Makefile.ref) and, independently, by central
 * differences of V and by the permutation symmetry of the surface
 * (tests/test_oracle_pes.py).
 */
#include <math.h>
#include "oracle.h"

#define SP(x) ((double)(float)(x))

/* H2 singlet curve of Schwenke (VH2OPT95, src/egrad_h3.f:322-417). */
static void h2_singlet(double r, double *e, double *de)
{
    /* A0..A11 are default-REAL literals, A12..A16 carry a D exponent. */
    const double a0 = SP(0.03537359271649620), a1 = SP(2.013977588700072),
                 a2 = SP(-2.827452449964767), a3 = SP(2.713257715593500),
                 a4 = SP(-2.792039234205731), a5 = SP(2.166542078766724),
                 a6 = SP(-1.272679684173909), a7 = SP(0.5630423099212294),
                 a8 = SP(-0.1879397372273814), a9 = SP(0.04719891893374140),
                 a10 = SP(-0.008851622656489644), a11 = SP(0.001224998776243630),
                 a12 = -1.227820520228028e-04, a13 = 8.638783190083473e-06,
                 a14 = -4.036967926499151e-07, a15 = 1.123286608335365e-08,
                 a16 = -1.406619156782167e-10;
    const double r0 = 3.5284882, dd = 0.160979391, c6 = 6.499027,
                 c8 = 124.3991, c10 = 3285.828;
    double p[16];
    p[1] = r;
    for (int i = 2; i < 16; ++i) p[i] = p[i - 1] * r;      /* r^i, i = 2..15 */
    const double r02 = r0 * r0, r04 = r02 * r02, r06 = r04 * r02;
    const double rr2 = p[2] + r02, rr4 = p[4] + r04, rr6 = p[6] + r06;
    const double rr25 = rr2 * rr2 * rr2 * rr2 * rr2;
    const double alpha = a0 / r + a1 + a2 * r + a3 * p[2] + a4 * p[3] + a5 * p[4]
                         + a6 * p[5] + a7 * p[6] + a8 * p[7] + a9 * p[8] + a10 * p[9]
                         + a11 * p[10] + a12 * p[11] + a13 * p[12] + a14 * p[13]
                         + a15 * p[14] + a16 * p[15];
    const double ex = exp(alpha);
    const double vsr = dd * (ex - 1.0) * (ex - 1.0) - dd;
    const double vlr = -c6 / rr6 - c8 / (rr4 * rr4) - c10 / rr25;
    *e = vsr + vlr;
    const double rr26 = rr25 * rr2, rr43 = rr4 * rr4 * rr4;
    const double dalpha = -a0 / p[2] + a2 + 2.0 * a3 * r + 3.0 * a4 * p[2]
                          + 4.0 * a5 * p[3] + 5.0 * a6 * p[4] + 6.0 * a7 * p[5]
                          + 7.0 * a8 * p[6] + 8.0 * a9 * p[7] + 9.0 * a10 * p[8]
                          + 10.0 * a11 * p[9] + 11.0 * a12 * p[10] + 12.0 * a13 * p[11]
                          + 13.0 * a14 * p[12] + 14.0 * a15 * p[13] + 15.0 * a16 * p[14];
    const double dvsr = 2.0 * dd * (ex - 1.0) * ex * dalpha;
    const double dvlr = 6.0 * c6 * p[5] / (rr6 * rr6) + 8.0 * c8 * p[3] / rr43
                        + 10.0 * c10 * r / rr26;
    *de = dvsr + dvlr;
}

/* H2 triplet curve: modified Truhlar-Horowitz form with the Johnson short-range
 * correction and a cubic transition (TRIPLET95, src/egrad_h3.f:251-320). */
static void h2_triplet(double r, double *e, double *de)
{
    const double rl = 0.95, rr = 1.15;
    const double a1 = SP(-0.0298546962), a2 = SP(-23.9604445036),
                 a3 = SP(-42.5185569474), a4 = SP(2.0382390988),
                 a5 = SP(-11.5214861455), a6 = SP(1.5309487826),
                 c1 = SP(-0.4106358351531854), c2 = SP(-0.0770355790707090),
                 c3 = SP(0.4303193846943223);
    if (r >= rr) {
        const double exdr = exp(-a4 * r);
        const double rsq = r * r;
        const double ra6 = pow(r, -a6);
        *e = a1 * (a2 + r + a3 * rsq + a5 * ra6) * exdr;
        const double ra61 = pow(r, -a6 - 1.0);
        *de = a1 * exdr * (1.0 - a2 * a4 + (2.0 * a3 - a4) * r - a3 * a4 * rsq
                           - a5 * a6 * ra61 - a4 * a5 * ra6);
    } else {
        const double dr = r - rl;
        double e1, de1;
        h2_singlet(r, &e1, &de1);
        if (r <= rl) {
            *e = e1 + c2 * dr + c3;
            *de = de1 + c2;
        } else {
            *e = e1 + c1 * dr * dr * dr + c2 * dr + c3;
            *de = de1 + 3.0 * c1 * dr * dr + c2;
        }
    }
}

/* London (LEPS-like) term with the eps^2 cusp rounding (H3LOND95, :419-480). */
static void london(const double r[3], double *v, double dv[3])
{
    double e1[3], de1[3], e3[3], de3[3], qq[3], jj[3];
    for (int i = 0; i < 3; ++i) {
        h2_singlet(r[i], &e1[i], &de1[i]);
        h2_triplet(r[i], &e3[i], &de3[i]);
        qq[i] = 0.5 * (e1[i] + e3[i]);
        jj[i] = 0.5 * (e1[i] - e3[i]);
    }
    const double sumq = qq[0] + qq[1] + qq[2];
    const double d21 = fabs(jj[1] - jj[0]), d32 = fabs(jj[2] - jj[1]), d31 = fabs(jj[2] - jj[0]);
    const double sumj = d21 * d21 + d32 * d32 + d31 * d31;
    const double jt = 0.5 * sumj + 1.0e-12;
    const double root = sqrt(jt);
    *v = sumq - root;
    dv[0] = 0.5 * (de1[0] + de3[0]) - 0.25 * (2.0 * jj[0] - jj[1] - jj[2]) * (de1[0] - de3[0]) / root;
    dv[1] = 0.5 * (de1[1] + de3[1]) - 0.25 * (2.0 * jj[1] - jj[2] - jj[0]) * (de1[1] - de3[1]) / root;
    dv[2] = 0.5 * (de1[2] + de3[2]) - 0.25 * (2.0 * jj[2] - jj[0] - jj[1]) * (de1[2] - de3[2]) / root;
}

/* |(r1-r2)(r2-r3)(r3-r1)| and its gradient (ACALC95, :549-574). */
static void asym_a(double r1, double r2, double r3, double *a, double da[3])
{
    double av = (r1 - r2) * (r2 - r3) * (r3 - r1);
    da[0] = (-2.0 * r1 + r2 + r3) * (r2 - r3);
    da[1] = (-2.0 * r2 + r3 + r1) * (r3 - r1);
    da[2] = (-2.0 * r3 + r1 + r2) * (r1 - r2);
    if (av < 0.0) {
        av = -av;
        da[0] = -da[0];
        da[1] = -da[1];
        da[2] = -da[2];
    }
    *a = av;
}

/* Asymmetric correction (VASCAL95, :482-547). */
static void vascal(const double rp[3], double *vas, double dvas[3])
{
    const double aa1 = SP(0.3788951192E-02), aa2 = SP(0.1478100901E-02),
                 aa3 = SP(-.1848513849E-03), aa4 = SP(0.9230803609E-05),
                 aa5 = SP(-.1293180255E-06), aa6 = SP(0.5237179303E+00),
                 aa7 = SP(-.1112326215E-02);
    const double r = rp[0] + rp[1] + rp[2];
    const double rsq = r * r, rcu = rsq * r;
    double a, da[3];
    asym_a(rp[0], rp[1], rp[2], &a, da);
    const double a2 = a * a, a3 = a2 * a, a4 = a3 * a, a5 = a4 * a;
    const double exp1 = exp(-aa1 * rcu);
    const double exp6 = exp(-aa6 * r);
    const double s = aa2 * a2 + aa3 * a3 + aa4 * a4 + aa5 * a5;
    *vas = s * exp1 + aa7 * a2 * exp6 / r;
    for (int i = 0; i < 3; ++i) {
        const double ds = (2.0 * aa2 * a + 3.0 * aa3 * a2 + 4.0 * aa4 * a3 + 5.0 * aa5 * a4) * da[i];
        dvas[i] = -3.0 * aa1 * rsq * s * exp1 + ds * exp1 - aa7 * a2 * exp6 / rsq
                  + 2.0 * aa7 * a * da[i] * exp6 / r - aa6 * aa7 * a2 * exp6 / r;
    }
}

/* Compactness switch T(R) for distances below 1.15 bohr (COMPAC95, :576-609). */
static int compact(const double r[3], double t[3], double dt[3])
{
    const double rr = 1.15, rp = 1.25;
    int n = 0;
    for (int i = 0; i < 3; ++i) {
        t[i] = 0.0;
        dt[i] = 0.0;
        if (r[i] < rr) ++n;
    }
    if (n == 0) return 0;
    for (int i = 0; i < 3; ++i) {
        if (r[i] < rr) {
            const double top = rr - r[i], bot = rp - r[i];
            const double top2 = top * top, top3 = top2 * top, bot2 = bot * bot;
            t[i] = top3 / bot;
            dt[i] = -3.0 * top2 / bot + top3 / bot2;
        }
    }
    return n;
}

/* Compact asymmetric correction (CASYM95, :611-705). */
static void casym(const double r[3], const double t[3], const double dt[3],
                  double *cas, double dcas[3])
{
    const double u1 = SP(0.2210243144E+00), u2 = SP(0.4367417579E+00), u3 = SP(0.6994985432E-02),
                 u4 = SP(0.1491096501E+01), u5 = SP(0.1602896673E+01), u6 = SP(-.2821747323E+01),
                 u7 = SP(0.4948310833E+00), u8 = SP(-.3540394679E-01), u9 = SP(-.3305809954E+01),
                 u10 = SP(0.3644382172E+01), u11 = SP(-.9997570970E+00), u12 = SP(0.7989919534E-01),
                 u13 = SP(-.1075807322E-02);
    double a, da[3];
    asym_a(r[0], r[1], r[2], &a, da);
    const double a2 = a * a;
    const double sumt = t[0] + t[1] + t[2];
    const double sr = r[0] + r[1] + r[2];
    const double pr = r[0] * r[1] * r[2];
    const double sr2 = sr * sr, sr3 = sr2 * sr;
    const double pr2 = pr * pr, pr3 = pr2 * pr;
    const double series = 1.0 + u4 / pr2 + u5 / pr + u6 + u7 * pr + u8 * pr2
                          + a * (u9 / pr2 + u10 / pr + u11 + u12 * pr + u13 * pr2);
    const double term1 = u1 / pow(pr, u2);
    const double eterm = exp(-u3 * sr3);
    *cas = sumt * a2 * term1 * series * eterm;
    const double dpr[3] = { r[1] * r[2], r[2] * r[0], r[0] * r[1] };
    for (int i = 0; i < 3; ++i) {
        const double dterm1 = -1.0 * u1 * u2 * pow(pr, -u2 - 1.0) * dpr[i];
        const double determ = eterm * (-3.0 * u3 * sr2);
        const double dseries = dpr[i] * (-2.0 * u4 / pr3 - u5 / pr2 + u7 + 2.0 * u8 * pr)
                               + da[i] * (u9 / pr2 + u10 / pr + u11 + u12 * pr + u13 * pr2)
                               + a * dpr[i] * (-2.0 * u9 / pr3 - u10 / pr2 + u12 + 2.0 * u13 * pr);
        dcas[i] = dt[i] * a2 * term1 * series * eterm
                  + 2.0 * a * da[i] * sumt * term1 * series * eterm
                  + dterm1 * sumt * a2 * series * eterm
                  + dseries * sumt * a2 * term1 * eterm
                  + determ * sumt * a2 * term1 * series;
    }
}

/* Correction for all compact geometries (CSYM95, :707-772). */
static void csym(const double r[3], double *cal, double dcal[3])
{
    const double rr = 1.15, rp = 1.25;
    const double v1 = SP(-.2071708868E+00), v2 = SP(-.5672350377E+00), v3 = SP(0.9058780367E-02);
    const double sr = r[0] + r[1] + r[2];
    const double sr2 = sr * sr, sr3 = sr * sr2;
    const double exp3 = exp(-v3 * sr3);
    const double dexp3 = -3.0 * v3 * sr2 * exp3;
    double g[3], dg[3];
    for (int i = 0; i < 3; ++i) {
        const double ri = r[i];
        const double a = rr - ri, a2 = a * a, a3 = a * a2;
        const double b = rp - ri, b2 = b * b;
        const double sumv = v1 + v1 * v2 * ri;
        g[i] = 0.0;
        dg[i] = 0.0;
        if (ri < rr) {
            g[i] = (a3 / b) * sumv;
            dg[i] = (a3 / b2) * sumv - 3.0 * (a2 / b) * sumv + (a3 / b) * v1 * v2;
        }
    }
    const double sumg = g[0] + g[1] + g[2];
    *cal = sumg * exp3;
    for (int i = 0; i < 3; ++i) dcal[i] = dg[i] * exp3 + sumg * dexp3;
}

/* One bending family (the VBEND A/B or CBEND A/B coefficient sets) shares this
 * coefficient layout; see VBCB95, src/egrad_h3.f:774-1430. */
typedef struct {
    double c11, c12, c13, c14, c15;
    double c21, c22, c23, c24;
    double c31, c32;
    double c41, c42, c43, c44;
    double c51, c52, c53;
    double c61, c62, c63;
    double c71, c72, c73, c74, c75;
    double c81, c82, c83, c84;
} bend_coef;

/* Bending terms VBNDA/VBNDB (always) and CBNDA/CBNDB (compact geometries). */
static void vbcb(const double rp[3], int icompc, const double t[3], const double dt[3],
                 double *vbnda, double *vbndb, double dvbnda[3], double dvbndb[3],
                 double *cbnda, double *cbndb, double dcbnda[3], double dcbndb[3])
{
    const double z58 = 0.625, z38 = 0.375;
    const double beta1 = 0.52, beta2 = 0.052, beta3 = 0.79;
    /* VBENDA (A..) and VBENDB (G..) coefficients, JUN21 fit. */
    const double va[16] = { SP(-.1838073394E+03), SP(0.1334593242E+02), SP(-.2358129537E+00),
                            SP(-.4668193478E+01), SP(0.7197506670E+01), SP(0.2162004275E+02),
                            SP(0.2106294028E+02), SP(0.4242962586E+01), SP(0.4453505045E+01),
                            SP(-.1456918088E+00), SP(-.1692657366E-01), SP(0.1279520698E+01),
                            SP(-.4898940075E+00), SP(0.1742295219E+03), SP(0.3142175348E+02),
                            SP(0.5152903406E+01) };
    const double vb[16] = { SP(-.4765732725E+02), SP(0.3648933563E+01), SP(-.7141145244E-01),
                            SP(0.1002349176E-01), SP(0.9989856329E-02), SP(-.4161953634E-02),
                            SP(0.9075807910E-03), SP(-.2693628729E+00), SP(-.1399065763E-01),
                            SP(-.1417634346E-01), SP(-.4870024792E-03), SP(0.1312231847E+00),
                            SP(-.4409850519E-01), SP(0.5382970863E+02), SP(0.4587102824E+01),
                            SP(0.1768550515E+01) };
    /* CBENDA (C..) and CBENDB (D..) coefficients, JUN21 fit. */
    const bend_coef ca = {
        SP(0.1860299931E+04), SP(-.6134458037E+03), SP(0.7337207161E+02), SP(-.2676717625E+04), SP(0.1344099415E+04),
        SP(0.1538913137E+03), SP(0.4348007369E+02), SP(0.1719720677E+03), SP(0.2115963042E+03),
        SP(-.7026089414E+02), SP(-.1300938992E+03),
        SP(0.1310273564E+01), SP(-.6175149574E+00), SP(-.2679089358E+02), SP(0.5577477171E+01),
        SP(-.3543353539E+04), SP(-.3740709591E+03), SP(0.7979303144E+02),
        SP(-.1104230585E+04), SP(0.4603572025E+04), SP(-.5593496634E+04),
        SP(-.1069406434E+02), SP(0.1021807153E+01), SP(0.6669828341E-01), SP(0.4168542348E+02), SP(0.1751608567E+02),
        SP(0.9486883238E+02), SP(-.1519334221E+02), SP(0.4024697252E+04), SP(-.2225159395E+02) };
    const bend_coef cb = {
        SP(0.4203543357E+03), SP(-.4922474096E+02), SP(0.3362942544E+00), SP(-.3827423082E+03), SP(0.1746726001E+03),
        SP(0.1699995737E-01), SP(0.1513036778E-01), SP(0.2659119354E-01), SP(-.5760387483E-02),
        SP(0.1020622621E+02), SP(0.1050536271E-01),
        SP(0.6836172780E+00), SP(-.1627858240E+00), SP(-.6925485045E+01), SP(0.1632567385E+01),
        SP(0.1083595009E+04), SP(0.4641431791E+01), SP(-.8233144461E+00),
        SP(-.6157225942E+02), SP(0.3094361471E+03), SP(-.3299631143E+03),
        SP(0.8866227120E+01), SP(-.1382126854E+01), SP(0.7620770145E-01), SP(-.5145757859E+02), SP(0.2046097265E+01),
        SP(0.2540775558E+01), SP(-.4889246569E+00), SP(-.1127439280E+04), SP(-.2269932295E+01) };

    const double r1 = rp[0], r2 = rp[1], r3 = rp[2];
    const double t1 = r1 * r1 - r2 * r2 - r3 * r3;
    const double t2 = r2 * r2 - r3 * r3 - r1 * r1;
    const double t3 = r3 * r3 - r1 * r1 - r2 * r2;
    /* cosines of the three internal angles */
    const double cs1 = t1 / (-2.0 * r2 * r3), cs2 = t2 / (-2.0 * r3 * r1), cs3 = t3 / (-2.0 * r1 * r2);
    const double sum = cs1 + cs2 + cs3;
    const double b1a = 1.0 - sum;
    const double cos3t1 = 4.0 * (cs1 * cs1 * cs1) - 3.0 * cs1;
    const double cos3t2 = 4.0 * (cs2 * cs2 * cs2) - 3.0 * cs2;
    const double cos3t3 = 4.0 * (cs3 * cs3 * cs3) - 3.0 * cs3;
    const double sumb = cos3t1 + cos3t2 + cos3t3;
    const double b1b = 1.0 - (z58 * sumb + z38 * sum);
    /* dC_j/dR_i */
    const double dc1dr1 = -r1 / (r2 * r3), dc2dr2 = -r2 / (r1 * r3), dc3dr3 = -r3 / (r1 * r2);
    const double dc1dr2 = (t1 / (r2 * r2) + 2.0) / (2.0 * r3);
    const double dc1dr3 = (t1 / (r3 * r3) + 2.0) / (2.0 * r2);
    const double dc2dr1 = (t2 / (r1 * r1) + 2.0) / (2.0 * r3);
    const double dc2dr3 = (t2 / (r3 * r3) + 2.0) / (2.0 * r1);
    const double dc3dr1 = (t3 / (r1 * r1) + 2.0) / (2.0 * r2);
    const double dc3dr2 = (t3 / (r2 * r2) + 2.0) / (2.0 * r1);
    double db1a[3], db1b[3];
    db1a[0] = -1.0 * (dc1dr1 + dc2dr1 + dc3dr1);
    db1a[1] = -1.0 * (dc1dr2 + dc2dr2 + dc3dr2);
    db1a[2] = -1.0 * (dc1dr3 + dc2dr3 + dc3dr3);
    const double d1 = 12.0 * cs1 * cs1 - 3.0, d2 = 12.0 * cs2 * cs2 - 3.0, d3 = 12.0 * cs3 * cs3 - 3.0;
    db1b[0] = -z58 * (d1 * dc1dr1 + d2 * dc2dr1 + d3 * dc3dr1) - z38 * (dc1dr1 + dc2dr1 + dc3dr1);
    db1b[1] = -z58 * (d1 * dc1dr2 + d2 * dc2dr2 + d3 * dc3dr2) - z38 * (dc1dr2 + dc2dr2 + dc3dr2);
    db1b[2] = -z58 * (d1 * dc1dr3 + d2 * dc2dr3 + d3 * dc3dr3) - z38 * (dc1dr3 + dc2dr3 + dc3dr3);

    const double r = r1 + r2 + r3, rsq = r * r;
    const double b2 = 1.0 / r1 + 1.0 / r2 + 1.0 / r3;
    const double b3 = (r2 - r1) * (r2 - r1) + (r3 - r2) * (r3 - r2) + (r1 - r3) * (r1 - r3);
    const double b3b = sqrt(b3 + 1.0e-12);
    const double db2[3] = { -1.0 / (r1 * r1), -1.0 / (r2 * r2), -1.0 / (r3 * r3) };
    double db3[6];
    db3[0] = 4.0 * r1 - 2.0 * r2 - 2.0 * r3;
    db3[1] = 4.0 * r2 - 2.0 * r3 - 2.0 * r1;
    db3[2] = 4.0 * r3 - 2.0 * r1 - 2.0 * r2;
    db3[3] = 0.5 * db3[0] / b3b;
    db3[4] = 0.5 * db3[1] / b3b;
    db3[5] = 0.5 * db3[2] / b3b;
    const double exp1 = exp(-beta1 * r), exp2 = exp(-beta2 * rsq), exp7 = exp(-beta3 * r);
    const double dexp1 = -beta1 * exp1, dexp2 = -2.0 * beta2 * r * exp2, dexp7 = -beta3 * exp7;

    /* VBEND A (coefficients va, angle function b1a) then B (vb, b1b). */
    for (int fam = 0; fam < 2; ++fam) {
        const double *c = fam == 0 ? va : vb;
        const double *db1 = fam == 0 ? db1a : db1b;
        const double b1 = fam == 0 ? b1a : b1b;
        const double b12 = b1 * b1, b13 = b12 * b1, b14 = b13 * b1, b15 = b14 * b1;
        const double asum = c[0] + c[1] * r + c[2] * rsq;
        const double bsum = c[3] * b12 + c[4] * b13 + c[5] * b14 + c[6] * b15;
        const double csum = c[7] * b1 * exp1 + c[8] * b12 * exp2;
        const double dsum1 = c[9] * exp1 + c[10] * exp2;
        const double dsum2 = c[11] * exp1 + c[12] * exp2;
        const double fsum = c[13] + c[14] * r + c[15] * rsq;
        const double v = b1 * asum * exp1 + bsum * exp2 + b2 * csum
                         + (b1 * b3 * dsum1 + b1 * b3b * dsum2) + b1 * fsum * exp7;
        const double dasum = c[1] + 2.0 * c[2] * r;
        const double dbsum = 2.0 * c[3] * b1 + 3.0 * c[4] * b12 + 4.0 * c[5] * b13 + 5.0 * c[6] * b14;
        const double ddsum1 = c[9] * dexp1 + c[10] * dexp2;
        const double ddsum2 = c[11] * dexp1 + c[12] * dexp2;
        const double dfsum = c[14] + 2.0 * c[15] * r;
        double *dv = fam == 0 ? dvbnda : dvbndb;
        for (int i = 0; i < 3; ++i) {
            const double dvb1 = db1[i] * asum * exp1 + b1 * dasum * exp1 + b1 * asum * dexp1;
            const double dvb2 = dbsum * db1[i] * exp2 + bsum * dexp2;
            const double dvb3 = db2[i] * csum + b2 * (c[7] * db1[i] * exp1 + c[7] * b1 * dexp1
                                                      + 2.0 * c[8] * b1 * db1[i] * exp2 + c[8] * b12 * dexp2);
            const double dvb4 = db1[i] * b3 * dsum1 + b1 * db3[i] * dsum1 + b1 * b3 * ddsum1
                                + db1[i] * b3b * dsum2 + b1 * db3[3 + i] * dsum2 + b1 * b3b * ddsum2;
            const double dvb5 = db1[i] * fsum * exp7 + b1 * dfsum * exp7 + b1 * fsum * dexp7;
            dv[i] = dvb1 + dvb2 + dvb3 + dvb4 + dvb5;
        }
        if (fam == 0) *vbnda = v; else *vbndb = v;
    }

    *cbnda = 0.0;
    *cbndb = 0.0;
    for (int i = 0; i < 3; ++i) { dcbnda[i] = 0.0; dcbndb[i] = 0.0; }
    if (icompc == 0) return;

    const double sumt = t[0] + t[1] + t[2];
    const double rcu = rsq * r;
    const double p = r1 * r2 * r3, psq = p * p, pcu = psq * p;
    const double dp[3] = { r2 * r3, r3 * r1, r1 * r2 };
    const double e7 = exp(-beta2 * rcu);
    const double de7 = -3.0 * beta2 * rsq * e7;
    for (int fam = 0; fam < 2; ++fam) {
        const bend_coef *c = fam == 0 ? &ca : &cb;
        const double *db1 = fam == 0 ? db1a : db1b;
        const double b1 = fam == 0 ? b1a : b1b;
        const double x1 = c->c51 + c->c83;     /* the refit moved C83 into C51 */
        const double b12 = b1 * b1, b13 = b12 * b1, b14 = b13 * b1, b15 = b14 * b1;
        const double asum = c->c11 + c->c12 * r + c->c13 * rsq + c->c14 / r + c->c15 / rsq;
        const double bsum = c->c21 * b12 + c->c22 * b13 + c->c23 * b14 + c->c24 * b15;
        const double csum = c->c31 * b1 * exp1 + c->c32 * b12 * exp2;
        const double dsum1 = c->c41 * exp1 + c->c42 * exp2;
        const double dsum2 = c->c43 * exp1 + c->c44 * exp2;
        const double fsum = x1 + c->c52 * r + c->c53 * rsq;
        const double gsum = c->c61 + c->c62 / r + c->c63 / rsq;
        const double aasum = c->c71 + c->c72 * p + c->c73 * psq + c->c74 / p + c->c75 / psq;
        const double ffsum = c->c81 * p + c->c82 * psq + c->c84 / psq;
        const double dasum = c->c12 + 2.0 * c->c13 * r - c->c14 / rsq - 2.0 * c->c15 / rcu;
        const double dbsum = 2.0 * c->c21 * b1 + 3.0 * c->c22 * b12 + 4.0 * c->c23 * b13 + 5.0 * c->c24 * b14;
        const double ddsum1 = c->c41 * dexp1 + c->c42 * dexp2;
        const double ddsum2 = c->c43 * dexp1 + c->c44 * dexp2;
        const double dfsum = c->c52 + 2.0 * c->c53 * r;
        const double dgsum = -c->c62 / rsq - 2.0 * c->c63 / rcu;
        const double daasum = c->c72 + 2.0 * c->c73 * p - c->c74 / psq - 2.0 * c->c75 / pcu;
        const double dffsum = c->c81 + 2.0 * c->c82 * p - 2.0 * c->c84 / pcu;
        const double cbnds = b1 * asum * exp1 / p + bsum * exp2 + b2 * csum
                             + (b1 * b3 * dsum1 + b1 * b3b * dsum2) + b1 * fsum * e7 / p
                             + b1 * gsum * e7 + b1 * aasum * exp2 + b1 * ffsum * e7;
        double *dc = fam == 0 ? dcbnda : dcbndb;
        for (int i = 0; i < 3; ++i) {
            const double dcb1 = db1[i] * asum * exp1 / p + b1 * dasum * exp1 / p
                                + b1 * asum * dexp1 / p - b1 * asum * exp1 * dp[i] / psq;
            const double dcb2 = dbsum * db1[i] * exp2 + bsum * dexp2;
            const double dcb3 = db2[i] * csum + b2 * (c->c31 * db1[i] * exp1 + c->c31 * b1 * dexp1
                                                      + c->c32 * 2.0 * b1 * db1[i] * exp2 + c->c32 * b12 * dexp2);
            const double dcb4 = db1[i] * b3 * dsum1 + b1 * db3[i] * dsum1 + b1 * b3 * ddsum1
                                + db1[i] * b3b * dsum2 + b1 * db3[3 + i] * dsum2 + b1 * b3b * ddsum2;
            const double dcb5 = db1[i] * fsum * e7 / p + b1 * dfsum * e7 / p
                                + b1 * fsum * de7 / p - b1 * fsum * e7 * dp[i] / psq;
            const double dcb6 = db1[i] * gsum * e7 + b1 * dgsum * e7 + b1 * gsum * de7;
            const double dcb7 = db1[i] * aasum * exp2 + b1 * daasum * dp[i] * exp2 + b1 * aasum * dexp2;
            const double dcb8 = db1[i] * ffsum * e7 + b1 * dffsum * dp[i] * e7 + b1 * ffsum * de7;
            const double dsum = dcb1 + dcb2 + dcb3 + dcb4 + dcb5 + dcb6 + dcb7 + dcb8;
            dc[i] = dt[i] * cbnds + sumt * dsum;
        }
        if (fam == 0) *cbnda = sumt * cbnds; else *cbndb = sumt * cbnds;
    }
}

void oracle_h3_pote(const double r[3], double *pe, double dpe[3])
{
    double vlon, dvlon[3], vas, dvas[3], t[3], dt[3];
    double cal = 0.0, cas = 0.0, dcal[3] = { 0, 0, 0 }, dcas[3] = { 0, 0, 0 };
    double vbnda, vbndb, dvbnda[3], dvbndb[3], cbnda, cbndb, dcbnda[3], dcbndb[3];
    london(r, &vlon, dvlon);
    vascal(r, &vas, dvas);
    const int icompc = compact(r, t, dt);
    if (icompc >= 1) {
        csym(r, &cal, dcal);
        casym(r, t, dt, &cas, dcas);
    }
    vbcb(r, icompc, t, dt, &vbnda, &vbndb, dvbnda, dvbndb, &cbnda, &cbndb, dcbnda, dcbndb);
    *pe = vlon + vas + vbnda + vbndb + cal + cas + cbnda + cbndb;
    for (int i = 0; i < 3; ++i)
        dpe[i] = dvlon[i] + dvas[i] + dvbnda[i] + dvbndb[i] + dcal[i] + dcas[i] + dcbnda[i] + dcbndb[i];
}

/* Bead-batched plugin entry, same contract as egrad_h3 (src/egrad_h3.f:29-77):
 * q(3,3,nbeads) in bohr, V(nbeads) in hartree, dVdq(3,3,nbeads) in Eh/bohr.
 * R(1)=|r_B-r_A|, R(2)=|r_A-r_C|, R(3)=|r_C-r_B|. */
void oracle_egrad_h3(const double *q, int natoms, int nbeads, double *v, double *dvdq)
{
    for (int k = 0; k < nbeads; ++k) {
        const double *x = q + (size_t)k * 3 * natoms;
        double *g = dvdq + (size_t)k * 3 * natoms;
        double ab[3], ac[3], bc[3], r[3], dr[3];
        for (int d = 0; d < 3; ++d) {
            ab[d] = x[3 + d] - x[d];
            ac[d] = x[d] - x[6 + d];
            bc[d] = x[6 + d] - x[3 + d];
        }
        r[0] = sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]);
        r[1] = sqrt(ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2]);
        r[2] = sqrt(bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2]);
        oracle_h3_pote(r, &v[k], dr);
        for (int d = 0; d < 3; ++d) {
            g[d] = dr[1] * ac[d] / r[1] - dr[0] * ab[d] / r[0];
            g[3 + d] = dr[0] * ab[d] / r[0] - dr[2] * bc[d] / r[2];
            g[6 + d] = dr[2] * bc[d] / r[2] - dr[1] * ac[d] / r[1];
        }
    }
}
