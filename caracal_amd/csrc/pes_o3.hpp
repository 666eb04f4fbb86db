// O + O2 surface of Z. Varga, Y. Paukku and D. G. Truhlar (J. Chem. Phys.
// 147, 154312 (2017), the 1 1A'' singlet) in the POTLIB form the reference
// evaluates (egrad_o3 / pot_o3 / o3pes, src/egrad_o3.f:33-185; its diatomic
// curve ev2gm2_o3, :503-565; D3(BJ) dispersion d3disp_o3 / edisp_o3,
// :906-1030; polynomial basis EvMorse_o3 .. evbas_o3, :323-501; constants of
// BLOCK DATA prmt_o3, :1057-1176).  One lane per bead, FP64.
//
// Inside: angstrom and kcal/mol.  R1 = O1-O2, R2 = O1-O3, R3 = O2-O3;
// v = 240.486 + sum_i V2(R_i) + sum_i D(R_i) + sum_k C_k B_k(m1, m2, m3),
// m_i = exp(-(R_i - ra)/a - (R_i - rb)^2/ab), B the 56 kept polynomials of
// the recursion below.  E = v Econv + Eref hartree, dE/dx = dv/dx_A Gconv
// hartree/bohr (the reference's rounded factors).
//
// The polynomials are symmetric in (m1, m2, m3) (every one is built from
// m1 + m2 + m3, m1 m2 + m1 m3 + m2 m3 and m1 m2 m3), so the partial by m_i
// is the partial by the first argument with m_i moved there: the derivative
// takes three passes of the recursion carrying one dual component each
// (value and d/dm of the first argument) instead of one pass of three, which
// keeps at most 24 live (value, derivative) pairs in registers.
//
// The reference converts the common-block distances to bohr and back inside
// its dispersion routine, so its gradient sees R moved by up to an ulp; here
// every term is evaluated at R as computed (a difference at rounding level).
#pragma once

namespace cam {
namespace o3 {

constexpr double kMegA = 0.83, kMegAB = 3.70, kMegRa = 1.25, kMegRb = 1.13;
constexpr double kCconv = 0.52917721092, kEconv = 0.159360144e-2, kGconv = 0.843297564e-3, kEref = -0.19172848;
constexpr double kTwoDe = 240.486;
constexpr double kCoef[56] = {
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

// p[k] = p[a] p[b] - p[s0] - p[s1] - p[s2] (-1: none) for k >= 3, k != 4
// (EvPoly_o3, src/egrad_o3.f:366-447); p[0] = 1, p[1], p[2], p[4] from the
// monomials.
struct Rec {
    signed char a, b, s0, s1, s2;
};
constexpr Rec kRec[67] = {
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

// evbas_o3 (:448-501): B is p without the constant and two-body entries
__host__ __device__ constexpr bool dropped(int k)
{
    return k == 0 || k == 1 || k == 3 || k == 6 || k == 10 || k == 15 || k == 22 || k == 30 || k == 40 ||
           k == 52 || k == 66;
}

// sum_k C_k B_k(x, y, z) and its partial by x (the MEG variable of R3 in the
// reference's monomial order; any, by symmetry)
__device__ inline void poly(double x, double y, double z, double &v, double &dvdx)
{
    double pv[67], pd[67];
    pv[0] = 1.0;
    pd[0] = 0.0;
    pv[1] = x + y + z;
    pd[1] = 1.0;
    const double m6 = y * z;
    pv[2] = x * y + x * z + m6;
    pd[2] = y + z;
    pv[4] = x * m6;
    pd[4] = m6;
    double acc = 0.0, dacc = 0.0;
    int n = 0;
#pragma unroll
    for (int k = 2; k < 67; ++k) {
        if (k >= 3 && k != 4) {
            const Rec r = kRec[k];
            double tv = pv[r.a] * pv[r.b], td = pd[r.a] * pv[r.b] + pv[r.a] * pd[r.b];
            if (r.s0 >= 0) { tv -= pv[r.s0]; td -= pd[r.s0]; }
            if (r.s1 >= 0) { tv -= pv[r.s1]; td -= pd[r.s1]; }
            if (r.s2 >= 0) { tv -= pv[r.s2]; td -= pd[r.s2]; }
            pv[k] = tv;
            pd[k] = td;
        }
        if (!dropped(k)) {
            acc += kCoef[n] * pv[k];
            dacc += kCoef[n] * pd[k];
            ++n;
        }
    }
    v = acc;
    dvdx = dacc;
}

// diatomic curve (imol 4) in kcal/mol and its r-derivative, r in angstrom
__device__ inline double v2(double r, double &dv)
{
    constexpr double alpha = 9.439784362354936e-1, beta = 1.262242998506810e0;
    constexpr double a[8] = {-1.488979427684798e3, 1.881435846488955e4, -1.053475425838226e5, 2.755135591229064e5,
                             -4.277588997761775e5, 4.404104009614092e5, -2.946204062950765e5, 1.176861219078620e5};
    double s = 0.0, g = 0.0, bk = alpha;
    const double r2 = r * r;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const double e = a[k] * exp(-bk * r2);
        s += e;
        g -= 2.0 * bk * r * e;
        bk *= beta;
    }
    dv = g * 627.509523475149e-3;
    return s * 627.509523475149e-3;
}

// D3(BJ) dispersion of one O-O pair at r (angstrom), kcal/mol, and its r-derivative
__device__ inline double disp(double r, double &de)
{
    constexpr double autoang = 0.52917726, autokcal = 627.509541;
    constexpr double s6 = 1.0, s8 = 2.0, a1 = 0.5299, a2 = 2.20, c6 = 12.8;
    constexpr double r2r4 = (double)2.59361680f;    // a default-REAL literal in the reference
    constexpr double c8 = 3.0 * c6 * r2r4 * r2r4;
    const double r0 = a1 * sqrt(c8 / c6) + a2;
    const double r02 = r0 * r0, r04 = r02 * r02, r06 = r04 * r02, r08 = r04 * r04;
    const double rb = r * (1.0 / autoang);
    const double x2 = rb * rb, x4 = x2 * x2, x6 = x4 * x2, x8 = x4 * x4;
    const double i6 = 1.0 / (x6 + r06), i8 = 1.0 / (x8 + r08);
    const double e6 = c6 * i6, e8 = c8 * i8;
    const double e6dr = -6.0 * c6 * x4 * rb * i6 * i6, e8dr = -8.0 * c8 * x6 * rb * i8 * i8;
    de = (-s6 * e6dr - s8 * e8dr) * (autokcal / autoang);
    return (-s6 * e6 - s8 * e8) * autokcal;
}

// egrad_o3 for one bead: x = (O1, O2, O3) x (x, y, z) in bohr
__device__ inline void egrad_bead(const double *x, double &v, double *g)
{
    constexpr int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
    double d[3][3], r[3], m[3], dm[3], dvdr[3];
    double e = kTwoDe;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int c = 0; c < 3; ++c) d[i][c] = (x[3 * pa[i] + c] - x[3 * pb[i] + c]) * kCconv;
        r[i] = sqrt(d[i][0] * d[i][0] + d[i][1] * d[i][1] + d[i][2] * d[i][2]);
        const double u = r[i] - kMegRb;
        m[i] = exp(-(r[i] - kMegRa) / kMegA - u * u / kMegAB);
        dm[i] = (-2.0 * u / kMegAB - 1.0 / kMegA) * m[i];
        double dv2, dd;
        e += v2(r[i], dv2);
        e += disp(r[i], dd);
        dvdr[i] = dv2 + dd;
    }
    // the reference's monomial order: rm1 = m(R3), rm2 = m(R2), rm3 = m(R1)
    double vp, d3, d2, d1, unused;
    poly(m[2], m[1], m[0], vp, d3);
    poly(m[1], m[2], m[0], unused, d2);
    poly(m[0], m[2], m[1], unused, d1);
    e += vp;
    dvdr[0] += d1 * dm[0];
    dvdr[1] += d2 * dm[1];
    dvdr[2] += d3 * dm[2];
    v = e * kEconv + kEref;
#pragma unroll
    for (int k = 0; k < 9; ++k) g[k] = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double f = dvdr[i] * kGconv / r[i];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            g[3 * pa[i] + c] += f * d[i][c];
            g[3 * pb[i] + c] -= f * d[i][c];
        }
    }
}

}  // namespace o3
}  // namespace cam
