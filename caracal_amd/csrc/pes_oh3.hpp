// OH + H2 surface of Schatz and Elgersma (Chem. Phys. Lett. 73, 21 (1980)) in
// the POTLIB form the reference evaluates (egrad_oh3, src/egrad_oh3.f:33-190;
// pot_oh3 and its three-body, four-body and H2O routines, :249-598; the
// constants of BLOCK DATA PTPACM_oh3, :620-633, all written with a D
// exponent).  One lane per bead, FP64.
//
// Atoms O, H1 (hydroxyl), H2, H3; distances R1 = O-H1, R2 = O-H2, R3 = O-H3,
// R4 = H1-H2, R5 = H1-H3, R6 = H2-H3.  V = three Morse pairs + LEPS over
// (R2, R3, R6) + H2O terms over (R1, R2, R4) and (R1, R3, R5) + a four-body
// Gaussian over (R2, R3, R4, R5) - 2 DE2 + DE3.
//
// The reference's routines share one derivative array: where an H2O switching
// factor is exactly zero (0.5 gamma s >= 43, an O-H distance beyond 37.6 bohr)
// the H2O routine leaves that entry unset and the previous routine's value
// stands in its place (src/egrad_oh3.f:590-594).  `shared` below carries that
// array so the device gives the reference's forces there too.
#pragma once

namespace cam {
namespace oh3 {

constexpr double kDe[3] = {0.148201, 0.0275690, 0.151548};
constexpr double kBeta[3] = {1.260580, 0.924180, 1.068620};
constexpr double kRe[3] = {1.863300, 2.907700, 1.428600};
constexpr double kSato = 0.10;
constexpr double kGam[3] = {2.399700, 1.058350, 2.399700};
constexpr double kReOH = 1.808090, kReHH = 2.861590;
constexpr double kCon[7] = {-.0015920, 0.026963, 0.0014689, 0.080011, 0.085816, -0.063179, 0.101380};
constexpr double kAlp[4] = {4.773, 7.14, 2.938, 5.28};
constexpr double kLam[4] = {0.10, 0.10, 0.20, 0.03};
constexpr double kAcon[2] = {0.10, 0.009};

// LEPS over (r0, r1, r2): pair 1's Morse constants for the two O-H
// distances, pair 3's for H-H.
__device__ inline double leps(double r0, double r1, double r2, double *dedr)
{
    const double rr[3] = {r0, r1, r2};
    double x[3], exch[3], b[3], energy = 0.0;
    constexpr double z = kSato, zpo = 1.0 + z, op3z = 1.0 + 3.0 * z, top3z = 2.0 * op3z, zp3 = z + 3.0,
                     tzp3 = 2.0 * zp3;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int k = i == 2 ? 2 : 0;
        const double do4z = kDe[k] / 4.0 / zpo;
        b[i] = kBeta[k] * do4z * 2.0;
        x[i] = exp(-kBeta[k] * (rr[i] - kRe[k]));
        const double coul = do4z * (zp3 * x[i] - top3z) * x[i];
        exch[i] = do4z * (op3z * x[i] - tzp3) * x[i];
        energy = energy + coul;
    }
    const double d12 = exch[0] - exch[1], d23 = exch[1] - exch[2], d31 = exch[2] - exch[0];
    const double rad = sqrt(d12 * d12 + d23 * d23 + d31 * d31);
    const double s2 = sqrt(2.0);
    energy = energy - rad / s2;
    const double s = exch[0] + exch[1] + exch[2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        dedr[i] = b[i] * x[i] * ((3.0 * exch[i] - s) / s2 * (op3z * x[i] - zp3) / rad - zp3 * x[i] + op3z);
    return energy;
}

// H2O term over (r_OH, r_OH', r_HH'); dedr holds the shared array on entry.
__device__ inline double h2o(double r0, double r1, double r2, double *dedr)
{
    double s[3], q[3], dq[3];
    s[0] = r0 - kReOH;
    s[2] = r1 - kReOH;
    s[1] = r2 - kReHH;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double x = 0.5 * kGam[i] * s[i];
        if (x >= 43.0) {
            q[i] = 0.0;
            dq[i] = 0.0;
        } else {
            q[i] = x >= 15.0 ? 2.0 / (1.0 + exp(2.0 * x)) : 1.0 - tanh(x);
            const double ch = cosh(x);
            dq[i] = -0.5 * kGam[i] / (ch * ch);
        }
    }
    const double *c = kCon;
    const double p = c[0] + c[1] * (s[0] + s[2]) + c[2] * s[1] + 0.5 * c[3] * (s[0] * s[0] + s[2] * s[2]) +
                     0.5 * c[4] * s[1] * s[1] + c[5] * s[1] * (s[0] + s[2]) + c[6] * s[0] * s[2];
    const double e = q[0] * q[1] * q[2] * p;
    const double dp[3] = {c[1] + c[3] * s[0] + c[5] * s[1] + c[6] * s[2], c[2] + c[4] * s[1] + c[5] * (s[0] + s[2]),
                          c[1] + c[3] * s[2] + c[5] * s[1] + c[6] * s[0]};
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (q[i] != 0.0) dedr[i] = e * (dq[i] / q[i] + (dp[i] / p));
    const double t = dedr[1];
    dedr[1] = dedr[2];
    dedr[2] = t;
    return e;
}

__device__ inline double four_body(const double *r, double *dedr)
{
    const double e1 = r[0] - kAlp[0], e2 = r[1] - kAlp[0], e3 = r[2] - kAlp[2], e4 = r[3] - kAlp[2];
    const double f1 = r[0] - kAlp[1], f2 = r[1] - kAlp[1], f3 = r[2] - kAlp[3], f4 = r[3] - kAlp[3];
    const double t1 =
        exp(-kLam[0] * (e1 * e1) - kLam[0] * (e2 * e2) - kLam[2] * (e3 * e3) - kLam[2] * (e4 * e4)) * kAcon[0];
    const double t2 =
        exp(-kLam[1] * (f1 * f1) - kLam[1] * (f2 * f2) - kLam[3] * (f3 * f3) - kLam[3] * (f4 * f4)) * kAcon[1];
    dedr[0] = -2.0 * (t1 * kLam[0] * e1 + t2 * kLam[1] * f1);
    dedr[1] = -2.0 * (t1 * kLam[0] * e2 + t2 * kLam[1] * f2);
    dedr[2] = -2.0 * (t1 * kLam[2] * e3 + t2 * kLam[3] * f3);
    dedr[3] = -2.0 * (t1 * kLam[2] * e4 + t2 * kLam[3] * f4);
    return t1 + t2;
}

// V and dV/dR of the six distances, summed in the reference's order.
__device__ inline double pot(const double *r, double *dvdr)
{
    auto morse = [](int k, double rr, double &dv) {
        const double ex = exp(-kBeta[k] * (rr - kRe[k]));
        const double u = 1.0 - ex;
        dv = 2.0 * kBeta[k] * kDe[k] * u * ex;
        return kDe[k] * (u * u);
    };
    double d0, d3, d4;
    double v = morse(0, r[0], d0) + morse(1, r[3], d3) + morse(1, r[4], d4);
    dvdr[0] = d0;
    dvdr[3] = d3;
    dvdr[4] = d4;
    double sh[4];
    v = v + leps(r[1], r[2], r[5], sh);
    dvdr[1] = sh[0];
    dvdr[2] = sh[1];
    dvdr[5] = sh[2];
    v = v + h2o(r[0], r[1], r[3], sh);
    dvdr[0] = dvdr[0] + sh[0];
    dvdr[1] = dvdr[1] + sh[1];
    dvdr[3] = dvdr[3] + sh[2];
    v = v + h2o(r[0], r[2], r[4], sh);
    dvdr[0] = dvdr[0] + sh[0];
    dvdr[2] = dvdr[2] + sh[1];
    dvdr[4] = dvdr[4] + sh[2];
    const double r4[4] = {r[1], r[2], r[3], r[4]};
    v = v + four_body(r4, sh);
    dvdr[1] = dvdr[1] + sh[0];
    dvdr[2] = dvdr[2] + sh[1];
    dvdr[3] = dvdr[3] + sh[2];
    dvdr[4] = dvdr[4] + sh[3];
    return v - 2.0 * kDe[1] + kDe[2];
}

// One bead: x = (O, H1, H2, H3) xyz, g the same layout.  The chain rule adds
// each distance's term in the reference's order (O: R1 R2 R3; H1: R1 R4 R5;
// H2: R2 R4 R6; H3: R3 R5 R6), with u = dV/dR * dx / R.
__device__ inline void egrad_bead(const double *x, double &v, double *g)
{
    constexpr int pa[6] = {1, 2, 3, 2, 3, 3}, pb[6] = {0, 0, 0, 1, 1, 2};
    double d[6][3], r[6], dvdr[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
#pragma unroll
        for (int c = 0; c < 3; ++c) d[m][c] = x[3 * pa[m] + c] - x[3 * pb[m] + c];
        r[m] = sqrt(d[m][0] * d[m][0] + d[m][1] * d[m][1] + d[m][2] * d[m][2]);
    }
    v = pot(r, dvdr);
    constexpr int terms[4][3] = {{0, 1, 2}, {0, 3, 4}, {1, 3, 5}, {2, 4, 5}};
#pragma unroll
    for (int at = 0; at < 4; ++at)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double acc = 0.0;
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                const int m = terms[at][t];
                const double u = dvdr[m] * d[m][c] / r[m];
                acc = pa[m] == at ? acc + u : acc - u;
            }
            g[3 * at + c] = acc;
        }
}

}  // namespace oh3
}  // namespace cam
