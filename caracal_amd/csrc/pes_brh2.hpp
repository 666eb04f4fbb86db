// BrH2 DIM-3C surface (diatomics-in-molecules plus a three-centre term;
// D. C. Clary, Chem. Phys. 71, 117 (1982); I. Last and M. Baer) in the POTLIB
// form the reference evaluates (egrad_brh2, src/egrad_brh2.f:33-79; POT_brh2,
// :207-441; VHX, :442-463; the constants of BLOCK DATA PTPARM, :466-483, all
// written with a D exponent).  One lane per bead, FP64.
//
// Atoms Ha, Br, Hc; distances R1 = Br-Ha, R2 = Br-Hc, R3 = Ha-Hc.  V = the
// lowest eigenvalue of a 4 x 4 DIM matrix (H2 singlet/triplet and the HBr 1S,
// 3S, 1P, 3P curves mixed by the two Br-H angles) + a three-centre overlap
// term + DHH, zero for Br far from H2(re).  The eigenvalue's derivatives are
// u^T (dH/dR) u with u its eigenvector (Hellmann-Feynman, :386-410).
//
// The reference finds all four eigenpairs with EISPACK (RSP, :485-1059).  Here
// the 4 x 4 matrix is diagonalised in registers by cyclic Jacobi rotations
// (every index compile-time, so nothing goes to scratch memory) until the
// off-diagonal elements sum to at most 2^-60 of the diagonal's magnitudes:
// the lowest eigenvalue and u^T D u do not depend on the method beyond
// rounding.  Where the lowest two eigenvalues lie close, the eigenvector (and
// with it the derivative) is ill-conditioned in either method; Jacobi's is
// the accurate one there (within 1e-14 of an 80-bit evaluation where
// EISPACK's derivatives are off by up to 6e-11).
#pragma once

namespace cam {
namespace brh2 {

constexpr double kEps = 1e-6, kRhh = 1.4016, kAhh = 1.0291, kDhh = 0.17447, kBhh = 0.018;
constexpr double kXih = 1.0, kXix = 1.6, kG = 0.22, kAlfw = 1.0;
constexpr double kRhx = 2.673, kAhx = 0.957, kDhx = 0.1439, kBhx = 0.012;
constexpr double kEtaHH = 0.393764, kEta3S = 0.322, kEta1P = 0.20286, kEta3P = 0.1771;
// initialize_brh2's derived constants (src/egrad_brh2.f:174-180); the
// normalisation of the H-Br overlap, 2 RHX (1 + t + t^2/3) exp(-t) at
// t = XIB RHX, is evaluated once per call (one exponential)
constexpr double kXib = 0.5 * (kXih + kXix), kC3 = 1.0 / 3.0;
constexpr double kGr = kG / kRhx;
constexpr double kRt34 = 0.25 * 1.7320508075688772, kRt38 = 0.5 * kRt34;

// HBr curves at r: (1S, 3S, 1P, 3P) and their r-derivatives
__device__ inline void vhx(double r, double v[4], double dv[4])
{
    const double rdif = r - kRhx, rdif2 = rdif * rdif;
    const double ex1 = exp(-kAhx * rdif), ex2 = exp(-kBhx * rdif2 * rdif);
    const double t1 = kDhx * ex1 * ex2, v1 = t1 * (ex1 + 2.0);
    v[0] = t1 * (ex1 - 2.0);
    v[1] = kEta3S * v1;
    v[2] = kEta1P * v1;
    v[3] = kEta3P * v1;
    const double a1 = 2.0 * kAhx * t1, a2 = 3.0 * kBhx * rdif2;
    const double dv1 = -a1 * (1.0 + ex1) - a2 * v1;
    dv[0] = a1 * (1.0 - ex1) - a2 * v[0];
    dv[1] = kEta3S * dv1;
    dv[2] = kEta1P * dv1;
    dv[3] = kEta3P * dv1;
}

// packed lower triangle, row-wise (0-based)
__host__ __device__ constexpr int pk(int i, int j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

// One Jacobi rotation in the (P, Q) plane of the packed symmetric matrix a,
// accumulated into the eigenvector columns v[.][P], v[.][Q].
template <int P, int Q>
__device__ inline void rotate(double a[10], double v[4][4])
{
    const double apq = a[pk(P, Q)];
    if (apq == 0.0) return;
    const double theta = (a[pk(Q, Q)] - a[pk(P, P)]) / (2.0 * apq);
    double t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
    t = theta < 0.0 ? -t : t;
    const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
    a[pk(P, P)] -= t * apq;
    a[pk(Q, Q)] += t * apq;
    a[pk(P, Q)] = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if (r != P && r != Q) {
            const double arp = a[pk(r, P)], arq = a[pk(r, Q)];
            a[pk(r, P)] = c * arp - s * arq;
            a[pk(r, Q)] = s * arp + c * arq;
        }
        const double vrp = v[r][P], vrq = v[r][Q];
        v[r][P] = c * vrp - s * vrq;
        v[r][Q] = s * vrp + c * vrq;
    }
}

// Lowest eigenpair of the symmetric 4 x 4 matrix a (packed, overwritten).
__device__ inline double lowest_eigenpair(double a[10], double u[4])
{
    double v[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) v[i][j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 64; ++sweep) {
        const double off = fabs(a[1]) + fabs(a[3]) + fabs(a[4]) + fabs(a[6]) + fabs(a[7]) + fabs(a[8]);
        const double diag = fabs(a[0]) + fabs(a[2]) + fabs(a[5]) + fabs(a[9]);
        if (off <= 0x1p-60 * diag) break;
        rotate<0, 1>(a, v);
        rotate<0, 2>(a, v);
        rotate<0, 3>(a, v);
        rotate<1, 2>(a, v);
        rotate<1, 3>(a, v);
        rotate<2, 3>(a, v);
    }
    double e = a[0];
#pragma unroll
    for (int i = 0; i < 4; ++i) u[i] = v[i][0];
#pragma unroll
    for (int m = 1; m < 4; ++m) {
        const double em = a[pk(m, m)];
        const bool lower = em < e;
        e = lower ? em : e;
#pragma unroll
        for (int i = 0; i < 4; ++i) u[i] = lower ? v[i][m] : u[i];
    }
    return e;
}

// POT_brh2: r = (R(1), R(2), R(3)) = (Br-Ha, Ha-Hc, Br-Hc), dedr in the same order
__device__ inline double pot(const double r[3], double dedr[3])
{
    double r1 = r[0], r2 = r[2], r3 = r[1];
    r1 = r1 > r2 + r3 ? r2 + r3 : r1;
    r2 = r2 > r1 + r3 ? r1 + r3 : r2;
    r3 = r3 > r1 + r2 ? r1 + r2 : r3;
    const double r1s = r1 * r1, r2s = r2 * r2, r3s = r3 * r3;
    const double r12 = r1 * r2, r13 = r1 * r3, r23 = r2 * r3;
    const double c3x = (r1s + r2s - r3s) * 0.5, c2x = (r1s + r3s - r2s) * 0.5, c1x = (r2s + r3s - r1s) * 0.5;

    // angle at Ha (sides R1, R3) and its derivatives
    double csa = c2x / r13;
    csa = fabs(csa) > 1.0 ? (csa > 0.0 ? 1.0 : -1.0) : csa;
    const double csa2 = csa * csa, sna2 = 1.0 - csa2;
    const bool lcol = sna2 < kEps;
    const double sna = sqrt(sna2);
    double t22 = 2.0 * csa, t11 = lcol ? 0.0 : 2.0 * (sna2 - csa2) / sna;
    const double sn2a = t22 * sna;
    double t = c3x / (r1 * r13);
    const double dcsa21 = t22 * t, dsn2a1 = t11 * t;
    t = -r2 / r13;
    const double dcsa22 = t22 * t, dsn2a2 = t11 * t;
    t = c1x / (r13 * r3);
    const double dcsa23 = t22 * t, dsn2a3 = t11 * t;

    // angle at Hc (sides R2, R3)
    double csb = c1x / r23;
    csb = fabs(csb) > 1.0 ? (csb > 0.0 ? 1.0 : -1.0) : csb;
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

    // cos^2 of the angle at Br
    t = c3x / r12;
    const double csg2 = t * t;
    t = 2.0 * t;
    const double dcsg21 = t * c2x / (r1 * r12), dcsg22 = t * c1x / (r12 * r2), dcsg23 = -t * r3 / r12;

    // H2 singlet and triplet
    double rdif = r3 - kRhh;
    double ex1 = exp(-kAhh * rdif);
    const double rdif2 = rdif * rdif, ex2 = exp(-kBhh * rdif * rdif2);
    const double h1 = kDhh * ex1 * ex2;
    const double v1hh = h1 * (ex1 - 2.0), v3hh = kEtaHH * h1 * (ex1 + 2.0);
    const double h2 = 2.0 * kAhh * h1, h3 = 3.0 * kBhh * rdif2;
    const double dv1hh = h2 * (1.0 - ex1) - h3 * v1hh;
    const double dv3hh = -h2 * kEtaHH * (1.0 + ex1) - h3 * v3hh;

    double w1[4], dw1[4], w2[4], dw2[4];
    vhx(r1, w1, dw1);
    vhx(r2, w2, dw2);
    const double s11 = w1[0] + 3.0 * w1[1], s12 = w2[0] + 3.0 * w2[1];
    const double s21 = 3.0 * w1[0] + w1[1], s22 = 3.0 * w2[0] + w2[1];
    const double s31 = w1[0] - w1[1], s32 = w2[0] - w2[1];
    const double p11 = w1[2] + 3.0 * w1[3], p12 = w2[2] + 3.0 * w2[3];
    const double p21 = 3.0 * w1[2] + w1[3], p22 = 3.0 * w2[2] + w2[3];
    const double p31 = w1[2] - w1[3], p32 = w2[2] - w2[3];
    const double u11 = s11 - p11, u12 = s12 - p12, u21 = s21 - p21, u22 = s22 - p22;
    const double u31 = s31 - p31, u32 = s32 - p32;

    // DIM matrix (src/egrad_brh2.f:322-341)
    double a[10];
    a[0] = v1hh + 0.25 * (s11 * csa2 + p11 * sna2 + s12 * csb2 + p12 * snb2);
    a[2] = v1hh + 0.25 * (s11 * sna2 + p11 * csa2 + s12 * snb2 + p12 * csb2);
    a[5] = v3hh + 0.25 * (s21 * csa2 + p21 * sna2 + s22 * csb2 + p22 * snb2);
    a[9] = v3hh + 0.25 * (s21 * sna2 + p21 * csa2 + s22 * snb2 + p22 * csb2);
    a[1] = lcol ? 0.0 : 0.125 * (u11 * sn2a - u12 * sn2b);
    a[3] = kRt34 * (s31 * csa2 + p31 * sna2 - s32 * csb2 - p32 * snb2);
    a[4] = lcol ? 0.0 : kRt38 * (u31 * sn2a + u32 * sn2b);
    a[6] = a[4];
    a[7] = kRt34 * (s31 * sna2 + p31 * csa2 - s32 * snb2 - p32 * csb2);
    a[8] = lcol ? 0.0 : 0.125 * (u21 * sn2a - u22 * sn2b);

    double u[4];
    const double e1 = lowest_eigenpair(a, u);
    // u_l u_k for the packed entries, off-diagonal ones twice
    double w[10];
#pragma unroll
    for (int l = 0; l < 4; ++l)
#pragma unroll
        for (int k = 0; k <= l; ++k) w[pk(l, k)] = (l == k ? 1.0 : 2.0) * u[l] * u[k];

    // u^T (dH/dR) u for R1, R2, R3 (the derivative matrices of :343-381)
    const double ds11 = dw1[0] + 3.0 * dw1[1], ds12 = dw2[0] + 3.0 * dw2[1];
    const double ds21 = 3.0 * dw1[0] + dw1[1], ds22 = 3.0 * dw2[0] + dw2[1];
    const double ds31 = dw1[0] - dw1[1], ds32 = dw2[0] - dw2[1];
    const double dp11 = dw1[2] + 3.0 * dw1[3], dp12 = dw2[2] + 3.0 * dw2[3];
    const double dp21 = 3.0 * dw1[2] + dw1[3], dp22 = 3.0 * dw2[2] + dw2[3];
    const double dp31 = dw1[2] - dw1[3], dp32 = dw2[2] - dw2[3];
    double d1, d2, d3;
    {
        const double ta = u11 * dcsa21 + u12 * dcsb21, tb = u21 * dcsa21 + u22 * dcsb21,
                     tc = u31 * dcsa21 - u32 * dcsb21;
        const double e5 = kRt38 * ((ds31 - dp31) * sn2a + u31 * dsn2a1 + u32 * dsn2b1);
        d1 = w[0] * 0.25 * (ds11 * csa2 + dp11 * sna2 + ta) +
             w[1] * 0.125 * ((ds11 - dp11) * sn2a + u11 * dsn2a1 - u12 * dsn2b1) +
             w[2] * 0.25 * (ds11 * sna2 + dp11 * csa2 - ta) + w[3] * kRt34 * (ds31 * csa2 + dp31 * sna2 + tc) +
             (w[4] + w[6]) * e5 + w[5] * 0.25 * (ds21 * csa2 + dp21 * sna2 + tb) +
             w[7] * kRt34 * (ds31 * sna2 + dp31 * csa2 - tc) +
             w[8] * 0.125 * ((ds21 - dp21) * sn2a + u21 * dsn2a1 - u22 * dsn2b1) +
             w[9] * 0.25 * (ds21 * sna2 + dp21 * csa2 - tb);
    }
    {
        const double ta = u11 * dcsa22 + u12 * dcsb22, tb = u21 * dcsa22 + u22 * dcsb22,
                     tc = u31 * dcsa22 - u32 * dcsb22;
        const double e5 = kRt38 * ((ds32 - dp32) * sn2b + u31 * dsn2a2 + u32 * dsn2b2);
        d2 = w[0] * 0.25 * (ds12 * csb2 + dp12 * snb2 + ta) +
             w[1] * 0.125 * (-(ds12 - dp12) * sn2b + u11 * dsn2a2 - u12 * dsn2b2) +
             w[2] * 0.25 * (ds12 * snb2 + dp12 * csb2 - ta) + w[3] * kRt34 * (-ds32 * csb2 - dp32 * snb2 + tc) +
             (w[4] + w[6]) * e5 + w[5] * 0.25 * (ds22 * csb2 + dp22 * snb2 + tb) +
             w[7] * kRt34 * (-ds32 * snb2 - dp32 * csb2 - tc) +
             w[8] * 0.125 * ((dp22 - ds22) * sn2b + u21 * dsn2a2 - u22 * dsn2b2) +
             w[9] * 0.25 * (ds22 * snb2 + dp22 * csb2 - tb);
    }
    {
        const double ta = 0.25 * (u11 * dcsa23 + u12 * dcsb23), tb = 0.25 * (u21 * dcsa23 + u22 * dcsb23),
                     tc = kRt34 * (u31 * dcsa23 - u32 * dcsb23);
        const double e5 = kRt38 * (u31 * dsn2a3 + u32 * dsn2b3);
        d3 = w[0] * (dv1hh + ta) + w[1] * 0.125 * (u11 * dsn2a3 - u12 * dsn2b3) + w[2] * (dv1hh - ta) + w[3] * tc +
             (w[4] + w[6]) * e5 + w[5] * (dv3hh + tb) - w[7] * tc + w[8] * 0.125 * (u21 * dsn2a3 - u22 * dsn2b3) +
             w[9] * (dv3hh - tb);
    }

    // three-centre term (src/egrad_brh2.f:413-432)
    const double tx = kXib * kRhx;
    const double inv_denom = 1.0 / (2.0 * kRhx * (1.0 + tx * (1.0 + kC3 * tx)) * exp(-tx));
    t = kXih * r3;
    ex1 = exp(-t);
    const double shh = (1.0 + t * (1.0 + kC3 * t)) * ex1;
    const double dshh = -kXih * t * (1.0 + t) * kC3 * ex1;
    t = kXib * r1;
    ex1 = exp(-t) * inv_denom;
    const double shx1 = r1 * (1.0 + t * (1.0 + kC3 * t)) * ex1;
    const double dshx1 = (1.0 + t * (1.0 - kC3 * t * t)) * ex1;
    t = kXib * r2;
    ex1 = exp(-t) * inv_denom;
    const double shx2 = r2 * (1.0 + t * (1.0 + kC3 * t)) * ex1;
    const double dshx2 = (1.0 + t * (1.0 - kC3 * t * t)) * ex1;
    rdif = r1 - r2;
    const double b1 = kGr * exp(-kAlfw * rdif * rdif);
    const double b2 = shh * (shx1 + shx2) + shx1 * shx2;
    const double b3 = 2.0 * kAlfw * rdif * csg2;
    dedr[0] = d1 + (dshx1 * (shh + shx2) * csg2 + b2 * (dcsg21 - b3)) * b1;
    dedr[2] = d2 + (dshx2 * (shh + shx1) * csg2 + b2 * (dcsg22 + b3)) * b1;
    dedr[1] = d3 + (dshh * (shx1 + shx2) * csg2 + b2 * dcsg23) * b1;
    return e1 + b2 * b1 * csg2 + kDhh;
}

// egrad_brh2 for one bead: x = (Ha, Br, Hc) x (x, y, z)
__device__ inline void egrad_bead(const double *x, double &v, double *g)
{
    double ab[3], ac[3], bc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        ab[c] = x[3 + c] - x[c];
        ac[c] = x[c] - x[6 + c];
        bc[c] = x[6 + c] - x[3 + c];
    }
    const double r[3] = {sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]),
                         sqrt(ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2]),
                         sqrt(bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2])};
    double d[3];
    v = pot(r, d);
    const double f0 = d[0] / r[0], f1 = d[1] / r[1], f2 = d[2] / r[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        g[c] = f1 * ac[c] - f0 * ab[c];
        g[3 + c] = f0 * ab[c] - f2 * bc[c];
        g[6 + c] = f2 * bc[c] - f1 * ac[c];
    }
}

}  // namespace brh2
}  // namespace cam
