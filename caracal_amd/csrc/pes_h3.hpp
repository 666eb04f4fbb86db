// BKMP2 H3 surface on the device (reference: src/egrad_h3.f:29-1476).
//
// One lane evaluates one bead.  The routine is written for registers: the
// reference's scratch arrays (VB, CB, DCB1..8, ...) become scalars that are
// folded into running sums, and the two bending families are one templated
// body.  T is the arithmetic type (double for the plain FP64 path, float for
// the mixed path); the caller chooses how results are accumulated.
// Coefficients that the reference writes as default-REAL literals are rounded
// to binary32 first, exactly as the compiled reference sees them.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>

#include "rt_math.hpp"

namespace cam {
namespace h3 {

#define CAM_SP(x) ((double)(float)(x))

template <typename T>
__device__ __forceinline__ void singlet(T r, T& e, T& de)
{
    // Schwenke H2 singlet (VH2OPT95): A0..A11 default REAL, A12..A16 REAL(8).
    const T a0 = T(CAM_SP(0.03537359271649620)), a1 = T(CAM_SP(2.013977588700072)),
            a2 = T(CAM_SP(-2.827452449964767)), a3 = T(CAM_SP(2.713257715593500)),
            a4 = T(CAM_SP(-2.792039234205731)), a5 = T(CAM_SP(2.166542078766724)),
            a6 = T(CAM_SP(-1.272679684173909)), a7 = T(CAM_SP(0.5630423099212294)),
            a8 = T(CAM_SP(-0.1879397372273814)), a9 = T(CAM_SP(0.04719891893374140)),
            a10 = T(CAM_SP(-0.008851622656489644)), a11 = T(CAM_SP(0.001224998776243630)),
            a12 = T(-1.227820520228028e-04), a13 = T(8.638783190083473e-06),
            a14 = T(-4.036967926499151e-07), a15 = T(1.123286608335365e-08),
            a16 = T(-1.406619156782167e-10);
    const T r0 = T(3.5284882), dd = T(0.160979391), c6 = T(6.499027), c8 = T(124.3991),
            c10 = T(3285.828);
    // Horner form of sum_{i=1..16} a_i r^(i-1) and of its derivative, always
    // in FP64: the terms reach |a_i r^(i-1)| ~ 1e3 around r = 4.5 bohr while
    // their sum stays O(1), so a binary32 evaluation would lose ~1e-4 of alpha.
    const double rd = double(r);
    double poly = double(a16);
    poly = poly * rd + double(a15); poly = poly * rd + double(a14); poly = poly * rd + double(a13);
    poly = poly * rd + double(a12); poly = poly * rd + double(a11); poly = poly * rd + double(a10);
    poly = poly * rd + double(a9);  poly = poly * rd + double(a8);  poly = poly * rd + double(a7);
    poly = poly * rd + double(a6);  poly = poly * rd + double(a5);  poly = poly * rd + double(a4);
    poly = poly * rd + double(a3);  poly = poly * rd + double(a2);  poly = poly * rd + double(a1);
    double dpoly = 15.0 * double(a16);
    dpoly = dpoly * rd + 14.0 * double(a15); dpoly = dpoly * rd + 13.0 * double(a14);
    dpoly = dpoly * rd + 12.0 * double(a13); dpoly = dpoly * rd + 11.0 * double(a12);
    dpoly = dpoly * rd + 10.0 * double(a11); dpoly = dpoly * rd + 9.0 * double(a10);
    dpoly = dpoly * rd + 8.0 * double(a9);   dpoly = dpoly * rd + 7.0 * double(a8);
    dpoly = dpoly * rd + 6.0 * double(a7);   dpoly = dpoly * rd + 5.0 * double(a6);
    dpoly = dpoly * rd + 4.0 * double(a5);   dpoly = dpoly * rd + 3.0 * double(a4);
    dpoly = dpoly * rd + 2.0 * double(a3);   dpoly = dpoly * rd + double(a2);
    const T rinv = T(1) / r;
    const T alpha = T(double(a0) / rd + poly);
    const T dalpha = T(-double(a0) / (rd * rd) + dpoly);
    const T ex = rt::exp(alpha);
    const T exm1 = ex - T(1);
    const T r2 = r * r, r4 = r2 * r2, r6 = r4 * r2;
    const T r02 = r0 * r0, r04 = r02 * r02, r06 = r04 * r02;
    const T rr2 = r2 + r02, rr4 = r4 + r04, rr6 = r6 + r06;
    const T i2 = T(1) / rr2, i4 = T(1) / rr4, i6 = T(1) / rr6;
    const T i25 = i2 * i2 * i2 * i2 * i2;
    e = dd * exm1 * exm1 - dd - c6 * i6 - c8 * i4 * i4 - c10 * i25;
    de = T(2) * dd * exm1 * ex * dalpha + T(6) * c6 * r4 * r * i6 * i6
         + T(8) * c8 * r2 * r * i4 * i4 * i4 + T(10) * c10 * r * i25 * i2;
}

template <typename T>
__device__ __forceinline__ void triplet(T r, T& e, T& de)
{
    // Modified Truhlar-Horowitz triplet with the Johnson short-range patch (TRIPLET95).
    const T rl = T(0.95), rr = T(1.15);
    const T a1 = T(CAM_SP(-0.0298546962)), a2 = T(CAM_SP(-23.9604445036)),
            a3 = T(CAM_SP(-42.5185569474)), a4 = T(CAM_SP(2.0382390988)),
            a5 = T(CAM_SP(-11.5214861455)), a6 = T(CAM_SP(1.5309487826)),
            c1 = T(CAM_SP(-0.4106358351531854)), c2 = T(CAM_SP(-0.0770355790707090)),
            c3 = T(CAM_SP(0.4303193846943223));
    if (r >= rr) {
        const T exdr = rt::exp(-a4 * r);
        const T ra6 = rt::exp(-a6 * rt::log(r));       // r^(-a6)
        const T rsq = r * r;
        e = a1 * (a2 + r + a3 * rsq + a5 * ra6) * exdr;
        de = a1 * exdr * (T(1) - a2 * a4 + (T(2) * a3 - a4) * r - a3 * a4 * rsq
                          - a5 * a6 * ra6 / r - a4 * a5 * ra6);
    } else {
        T e1, de1;
        singlet(r, e1, de1);
        const T dr = r - rl;
        if (r <= rl) {
            e = e1 + c2 * dr + c3;
            de = de1 + c2;
        } else {
            e = e1 + c1 * dr * dr * dr + c2 * dr + c3;
            de = de1 + T(3) * c1 * dr * dr + c2;
        }
    }
}

// |(r1-r2)(r2-r3)(r3-r1)| and gradient (ACALC95).
template <typename T>
__device__ __forceinline__ T asym_a(const T r[3], T da[3])
{
    T a = (r[0] - r[1]) * (r[1] - r[2]) * (r[2] - r[0]);
    da[0] = (-T(2) * r[0] + r[1] + r[2]) * (r[1] - r[2]);
    da[1] = (-T(2) * r[1] + r[2] + r[0]) * (r[2] - r[0]);
    da[2] = (-T(2) * r[2] + r[0] + r[1]) * (r[0] - r[1]);
    if (a < T(0)) {
        a = -a;
        da[0] = -da[0]; da[1] = -da[1]; da[2] = -da[2];
    }
    return a;
}

// Bending coefficients in the order the reference numbers them.
// VBENDA (A11..A53) and VBENDB (G11..G53), 16 values each.
static __constant__ const double kVbend[2][16] = {
        { CAM_SP(-.1838073394E+03), CAM_SP(0.1334593242E+02), CAM_SP(-.2358129537E+00),
          CAM_SP(-.4668193478E+01), CAM_SP(0.7197506670E+01), CAM_SP(0.2162004275E+02),
          CAM_SP(0.2106294028E+02), CAM_SP(0.4242962586E+01), CAM_SP(0.4453505045E+01),
          CAM_SP(-.1456918088E+00), CAM_SP(-.1692657366E-01), CAM_SP(0.1279520698E+01),
          CAM_SP(-.4898940075E+00), CAM_SP(0.1742295219E+03), CAM_SP(0.3142175348E+02),
          CAM_SP(0.5152903406E+01) },
        { CAM_SP(-.4765732725E+02), CAM_SP(0.3648933563E+01), CAM_SP(-.7141145244E-01),
          CAM_SP(0.1002349176E-01), CAM_SP(0.9989856329E-02), CAM_SP(-.4161953634E-02),
          CAM_SP(0.9075807910E-03), CAM_SP(-.2693628729E+00), CAM_SP(-.1399065763E-01),
          CAM_SP(-.1417634346E-01), CAM_SP(-.4870024792E-03), CAM_SP(0.1312231847E+00),
          CAM_SP(-.4409850519E-01), CAM_SP(0.5382970863E+02), CAM_SP(0.4587102824E+01),
          CAM_SP(0.1768550515E+01) } };

// CBENDA (C11..C15, C21..C24, C31..C32, C41..C44, C51..C53, C61..C63,
// C71..C75, C81..C84) and CBENDB (D..), 30 values each; the refit moves C83
// into C51, which pote applies.
static __constant__ const double kCbend[2][30] = {
        { CAM_SP(0.1860299931E+04), CAM_SP(-.6134458037E+03), CAM_SP(0.7337207161E+02),
          CAM_SP(-.2676717625E+04), CAM_SP(0.1344099415E+04), CAM_SP(0.1538913137E+03),
          CAM_SP(0.4348007369E+02), CAM_SP(0.1719720677E+03), CAM_SP(0.2115963042E+03),
          CAM_SP(-.7026089414E+02), CAM_SP(-.1300938992E+03), CAM_SP(0.1310273564E+01),
          CAM_SP(-.6175149574E+00), CAM_SP(-.2679089358E+02), CAM_SP(0.5577477171E+01),
          CAM_SP(-.3543353539E+04), CAM_SP(-.3740709591E+03), CAM_SP(0.7979303144E+02),
          CAM_SP(-.1104230585E+04), CAM_SP(0.4603572025E+04), CAM_SP(-.5593496634E+04),
          CAM_SP(-.1069406434E+02), CAM_SP(0.1021807153E+01), CAM_SP(0.6669828341E-01),
          CAM_SP(0.4168542348E+02), CAM_SP(0.1751608567E+02), CAM_SP(0.9486883238E+02),
          CAM_SP(-.1519334221E+02), CAM_SP(0.4024697252E+04), CAM_SP(-.2225159395E+02) },
        { CAM_SP(0.4203543357E+03), CAM_SP(-.4922474096E+02), CAM_SP(0.3362942544E+00),
          CAM_SP(-.3827423082E+03), CAM_SP(0.1746726001E+03), CAM_SP(0.1699995737E-01),
          CAM_SP(0.1513036778E-01), CAM_SP(0.2659119354E-01), CAM_SP(-.5760387483E-02),
          CAM_SP(0.1020622621E+02), CAM_SP(0.1050536271E-01), CAM_SP(0.6836172780E+00),
          CAM_SP(-.1627858240E+00), CAM_SP(-.6925485045E+01), CAM_SP(0.1632567385E+01),
          CAM_SP(0.1083595009E+04), CAM_SP(0.4641431791E+01), CAM_SP(-.8233144461E+00),
          CAM_SP(-.6157225942E+02), CAM_SP(0.3094361471E+03), CAM_SP(-.3299631143E+03),
          CAM_SP(0.8866227120E+01), CAM_SP(-.1382126854E+01), CAM_SP(0.7620770145E-01),
          CAM_SP(-.5145757859E+02), CAM_SP(0.2046097265E+01), CAM_SP(0.2540775558E+01),
          CAM_SP(-.4889246569E+00), CAM_SP(-.1127439280E+04), CAM_SP(-.2269932295E+01) } };


// Bending terms (VBCB95).  Evaluated in FP64 on both paths: the A- and
// F-sums of the VB families are O(100) and their exponentially damped products
// cancel to O(0.1) near the saddle, so binary32 exponentials there would cost
// ~1e-6 Eh/bohr in the gradient.  v and dv accumulate in place.
template <typename T>
__device__ __forceinline__ void bend(const T r[3], int icompc, const T t[3], const T dt[3], T& v, T dv[3])
{
    const T rs = r[0] + r[1] + r[2];
    const T rsq = rs * rs;
    const T r1 = r[0], r2 = r[1], r3 = r[2];
    const T t1 = r1 * r1 - r2 * r2 - r3 * r3;
    const T t2 = r2 * r2 - r3 * r3 - r1 * r1;
    const T t3 = r3 * r3 - r1 * r1 - r2 * r2;
    const T cs1 = t1 / (-T(2) * r2 * r3), cs2 = t2 / (-T(2) * r3 * r1), cs3 = t3 / (-T(2) * r1 * r2);
    const T csum = cs1 + cs2 + cs3;
    const T cos3 = (T(4) * cs1 * cs1 * cs1 - T(3) * cs1) + (T(4) * cs2 * cs2 * cs2 - T(3) * cs2)
                   + (T(4) * cs3 * cs3 * cs3 - T(3) * cs3);
    const T b1fam[2] = { T(1) - csum, T(1) - (T(0.625) * cos3 + T(0.375) * csum) };
    // dC_j/dR_i
    const T dc1[3] = { -r1 / (r2 * r3), (t1 / (r2 * r2) + T(2)) / (T(2) * r3), (t1 / (r3 * r3) + T(2)) / (T(2) * r2) };
    const T dc2[3] = { (t2 / (r1 * r1) + T(2)) / (T(2) * r3), -r2 / (r1 * r3), (t2 / (r3 * r3) + T(2)) / (T(2) * r1) };
    const T dc3[3] = { (t3 / (r1 * r1) + T(2)) / (T(2) * r2), (t3 / (r2 * r2) + T(2)) / (T(2) * r1), -r3 / (r1 * r2) };
    const T q1 = T(12) * cs1 * cs1 - T(3), q2 = T(12) * cs2 * cs2 - T(3), q3 = T(12) * cs3 * cs3 - T(3);
    T db1[2][3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        db1[0][i] = -(dc1[i] + dc2[i] + dc3[i]);
        db1[1][i] = -T(0.625) * (q1 * dc1[i] + q2 * dc2[i] + q3 * dc3[i]) - T(0.375) * (dc1[i] + dc2[i] + dc3[i]);
    }
    const T b2 = T(1) / r1 + T(1) / r2 + T(1) / r3;
    const T db2[3] = { -T(1) / (r1 * r1), -T(1) / (r2 * r2), -T(1) / (r3 * r3) };
    const T b3 = (r2 - r1) * (r2 - r1) + (r3 - r2) * (r3 - r2) + (r1 - r3) * (r1 - r3);
    const T b3b = rt::sqrt(b3 + T(1.0e-12));
    const T db3[3] = { T(4) * r1 - T(2) * r2 - T(2) * r3, T(4) * r2 - T(2) * r3 - T(2) * r1,
                       T(4) * r3 - T(2) * r1 - T(2) * r2 };
    const T ib3b = T(0.5) / b3b;
    const T beta1 = T(0.52), beta2 = T(0.052), beta3 = T(0.79);
    const T exp1 = rt::exp(-beta1 * rs), exp2 = rt::exp(-beta2 * rsq), exp7 = rt::exp(-beta3 * rs);
    const T dexp1 = -beta1 * exp1, dexp2 = -T(2) * beta2 * rs * exp2, dexp7 = -beta3 * exp7;
#pragma unroll
    for (int fam = 0; fam < 2; ++fam) {
        const double* c = kVbend[fam];
        const T b1 = b1fam[fam];
        const T b12 = b1 * b1, b13 = b12 * b1, b14 = b13 * b1, b15 = b14 * b1;
        const T asum = T(c[0]) + T(c[1]) * rs + T(c[2]) * rsq;
        const T bsum = T(c[3]) * b12 + T(c[4]) * b13 + T(c[5]) * b14 + T(c[6]) * b15;
        const T csm = T(c[7]) * b1 * exp1 + T(c[8]) * b12 * exp2;
        const T ds1 = T(c[9]) * exp1 + T(c[10]) * exp2;
        const T ds2 = T(c[11]) * exp1 + T(c[12]) * exp2;
        const T fsum = T(c[13]) + T(c[14]) * rs + T(c[15]) * rsq;
        v += b1 * asum * exp1 + bsum * exp2 + b2 * csm + (b1 * b3 * ds1 + b1 * b3b * ds2) + b1 * fsum * exp7;
        const T dasum = T(c[1]) + T(2) * T(c[2]) * rs;
        const T dbsum = T(2) * T(c[3]) * b1 + T(3) * T(c[4]) * b12 + T(4) * T(c[5]) * b13 + T(5) * T(c[6]) * b14;
        const T dds1 = T(c[9]) * dexp1 + T(c[10]) * dexp2;
        const T dds2 = T(c[11]) * dexp1 + T(c[12]) * dexp2;
        const T dfsum = T(c[14]) + T(2) * T(c[15]) * rs;
        // parts of the derivative that do not depend on i
        const T k0 = b1 * dasum * exp1 + b1 * asum * dexp1 + bsum * dexp2
                     + b2 * (T(c[7]) * b1 * dexp1 + T(c[8]) * b12 * dexp2)
                     + b1 * b3 * dds1 + b1 * b3b * dds2 + b1 * dfsum * exp7 + b1 * fsum * dexp7;
        const T kb1 = asum * exp1 + dbsum * exp2 + b2 * (T(c[7]) * exp1 + T(2) * T(c[8]) * b1 * exp2)
                      + b3 * ds1 + b3b * ds2 + fsum * exp7;
#pragma unroll
        for (int i = 0; i < 3; ++i)
            dv[i] += k0 + db1[fam][i] * kb1 + db2[i] * csm + b1 * db3[i] * ds1 + b1 * db3[i] * ib3b * ds2;
    }
    if (icompc == 0) return;
    const T sumt = t[0] + t[1] + t[2];
    const T rcu = rsq * rs;
    const T p = r1 * r2 * r3;
    const T ip = T(1) / p;
    const T ip2 = ip * ip, ip3 = ip2 * ip;
    const T dp[3] = { r2 * r3, r3 * r1, r1 * r2 };
    const T e7 = rt::exp(-beta2 * rcu);
    const T de7 = -T(3) * beta2 * rsq * e7;
    const T irs = T(1) / rs, irs2 = irs * irs, irs3 = irs2 * irs;
#pragma unroll
    for (int fam = 0; fam < 2; ++fam) {
        const double* c = kCbend[fam];
        const T b1 = b1fam[fam];
        const T b12 = b1 * b1, b13 = b12 * b1, b14 = b13 * b1, b15 = b14 * b1;
        const T x1 = T(c[15] + c[28]);          // C51 + C83
        const T asum = T(c[0]) + T(c[1]) * rs + T(c[2]) * rsq + T(c[3]) * irs + T(c[4]) * irs2;
        const T bsum = T(c[5]) * b12 + T(c[6]) * b13 + T(c[7]) * b14 + T(c[8]) * b15;
        const T csm = T(c[9]) * b1 * exp1 + T(c[10]) * b12 * exp2;
        const T ds1 = T(c[11]) * exp1 + T(c[12]) * exp2;
        const T ds2 = T(c[13]) * exp1 + T(c[14]) * exp2;
        const T fsum = x1 + T(c[16]) * rs + T(c[17]) * rsq;
        const T gsum = T(c[18]) + T(c[19]) * irs + T(c[20]) * irs2;
        const T aasum = T(c[21]) + T(c[22]) * p + T(c[23]) * p * p + T(c[24]) * ip + T(c[25]) * ip2;
        const T ffsum = T(c[26]) * p + T(c[27]) * p * p + T(c[29]) * ip2;
        const T dasum = T(c[1]) + T(2) * T(c[2]) * rs - T(c[3]) * irs2 - T(2) * T(c[4]) * irs3;
        const T dbsum = T(2) * T(c[5]) * b1 + T(3) * T(c[6]) * b12 + T(4) * T(c[7]) * b13 + T(5) * T(c[8]) * b14;
        const T dds1 = T(c[11]) * dexp1 + T(c[12]) * dexp2;
        const T dds2 = T(c[13]) * dexp1 + T(c[14]) * dexp2;
        const T dfsum = T(c[16]) + T(2) * T(c[17]) * rs;
        const T dgsum = -T(c[19]) * irs2 - T(2) * T(c[20]) * irs3;
        const T daasum = T(c[22]) + T(2) * T(c[23]) * p - T(c[24]) * ip2 - T(2) * T(c[25]) * ip3;
        const T dffsum = T(c[26]) + T(2) * T(c[27]) * p - T(2) * T(c[29]) * ip3;
        const T cbnds = b1 * asum * exp1 * ip + bsum * exp2 + b2 * csm + (b1 * b3 * ds1 + b1 * b3b * ds2)
                        + b1 * fsum * e7 * ip + b1 * gsum * e7 + b1 * aasum * exp2 + b1 * ffsum * e7;
        v += sumt * cbnds;
        const T k0 = b1 * dasum * exp1 * ip + b1 * asum * dexp1 * ip + bsum * dexp2
                     + b2 * (T(c[9]) * b1 * dexp1 + T(c[10]) * b12 * dexp2)
                     + b1 * b3 * dds1 + b1 * b3b * dds2
                     + b1 * dfsum * e7 * ip + b1 * fsum * de7 * ip
                     + b1 * dgsum * e7 + b1 * gsum * de7 + b1 * aasum * dexp2 + b1 * ffsum * de7;
        const T kb1 = asum * exp1 * ip + dbsum * exp2 + b2 * (T(c[9]) * exp1 + T(2) * T(c[10]) * b1 * exp2)
                      + b3 * ds1 + b3b * ds2 + fsum * e7 * ip + gsum * e7 + aasum * exp2 + ffsum * e7;
        const T kdp = -b1 * asum * exp1 * ip2 - b1 * fsum * e7 * ip2 + b1 * daasum * exp2 + b1 * dffsum * e7;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const T dsum = k0 + db1[fam][i] * kb1 + db2[i] * csm + b1 * db3[i] * ds1
                           + b1 * db3[i] * ib3b * ds2 + dp[i] * kdp;
            dv[i] += dt[i] * cbnds + sumt * dsum;
        }
    }
}

// London term (H3LOND95): singlet and triplet H2 curves combined.
template <typename T>
__device__ __forceinline__ void london(const T r[3], T& v, T dv[3])
{
    T e1[3], de1[3], e3[3], de3[3], jj[3];
    T sumq = T(0);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        singlet(r[i], e1[i], de1[i]);
        triplet(r[i], e3[i], de3[i]);
        sumq += T(0.5) * (e1[i] + e3[i]);
        jj[i] = T(0.5) * (e1[i] - e3[i]);
    }
    const T d21 = jj[1] - jj[0], d32 = jj[2] - jj[1], d31 = jj[2] - jj[0];
    const T root = rt::sqrt(T(0.5) * (d21 * d21 + d32 * d32 + d31 * d31) + T(1.0e-12));
    v = sumq - root;
    const T q = T(0.25) / root;
    dv[0] = T(0.5) * (de1[0] + de3[0]) - (T(2) * jj[0] - jj[1] - jj[2]) * (de1[0] - de3[0]) * q;
    dv[1] = T(0.5) * (de1[1] + de3[1]) - (T(2) * jj[1] - jj[2] - jj[0]) * (de1[1] - de3[1]) * q;
    dv[2] = T(0.5) * (de1[2] + de3[2]) - (T(2) * jj[2] - jj[0] - jj[1]) * (de1[2] - de3[2]) * q;
}

// V(R1,R2,R3) and dV/dR_i of the BKMP2 surface (POTE).  R(1)=r_AB, R(2)=r_AC, R(3)=r_BC.
// The London and bending terms are evaluated in FP64 for either T: their
// O(1) parts cancel to the O(1e-2) forces of the saddle region, and a
// binary32 evaluation (inputs included) there costs ~1e-5 of max|F|.  T is
// the type of the asymmetric and compact corrections.
template <typename T>
__device__ __forceinline__ void pote(const double rd[3], double& vd, double dvd[3])
{
    london<double>(rd, vd, dvd);
    const T r[3] = { T(rd[0]), T(rd[1]), T(rd[2]) };
    // FP64 keeps the reference's order of accumulation; binary32 corrections
    // are summed on their own and added once.
    constexpr bool F64 = std::is_same<T, double>::value;
    T v = F64 ? T(vd) : T(0), dv[3] = { F64 ? T(dvd[0]) : T(0), F64 ? T(dvd[1]) : T(0), F64 ? T(dvd[2]) : T(0) };
    const T rs = r[0] + r[1] + r[2];
    const T rsq = rs * rs;
    T da[3];
    const T a = asym_a(r, da);
    // ---- asymmetric correction (VASCAL95) ----
    {
        const T aa1 = T(CAM_SP(0.3788951192E-02)), aa2 = T(CAM_SP(0.1478100901E-02)),
                aa3 = T(CAM_SP(-.1848513849E-03)), aa4 = T(CAM_SP(0.9230803609E-05)),
                aa5 = T(CAM_SP(-.1293180255E-06)), aa6 = T(CAM_SP(0.5237179303E+00)),
                aa7 = T(CAM_SP(-.1112326215E-02));
        const T a2 = a * a, a3 = a2 * a, a4 = a3 * a, a5 = a4 * a;
        const T exp1 = rt::exp(-aa1 * rsq * rs);
        const T exp6 = rt::exp(-aa6 * rs);
        const T s = aa2 * a2 + aa3 * a3 + aa4 * a4 + aa5 * a5;
        const T ir = T(1) / rs;
        v += s * exp1 + aa7 * a2 * exp6 * ir;
        const T common = -T(3) * aa1 * rsq * s * exp1 - aa7 * a2 * exp6 * ir * ir - aa6 * aa7 * a2 * exp6 * ir;
        const T dsda = (T(2) * aa2 * a + T(3) * aa3 * a2 + T(4) * aa4 * a3 + T(5) * aa5 * a4);
#pragma unroll
        for (int i = 0; i < 3; ++i)
            dv[i] += common + dsda * da[i] * exp1 + T(2) * aa7 * a * da[i] * exp6 * ir;
    }
    // ---- compactness switch (COMPAC95) ----
    const T rrc = T(1.15), rpc = T(1.25);
    T t[3], dt[3];
    int icompc = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        t[i] = T(0);
        dt[i] = T(0);
        if (r[i] < rrc) {
            ++icompc;
            const T top = rrc - r[i], bot = rpc - r[i];
            const T top2 = top * top, top3 = top2 * top;
            t[i] = top3 / bot;
            dt[i] = -T(3) * top2 / bot + top3 / (bot * bot);
        }
    }
    if (icompc > 0) {
        // ---- compact symmetric (CSYM95) ----
        const T v1 = T(CAM_SP(-.2071708868E+00)), v2 = T(CAM_SP(-.5672350377E+00)),
                v3 = T(CAM_SP(0.9058780367E-02));
        const T exp3 = rt::exp(-v3 * rsq * rs);
        const T dexp3 = -T(3) * v3 * rsq * exp3;
        T g[3], dg[3], sumg = T(0);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            g[i] = T(0);
            dg[i] = T(0);
            if (r[i] < rrc) {
                const T x = rrc - r[i], y = rpc - r[i];
                const T sv = v1 + v1 * v2 * r[i];
                const T x3y = x * x * x / y;
                g[i] = x3y * sv;
                dg[i] = x3y / y * sv - T(3) * (x * x / y) * sv + x3y * v1 * v2;
            }
            sumg += g[i];
        }
        v += sumg * exp3;
#pragma unroll
        for (int i = 0; i < 3; ++i) dv[i] += dg[i] * exp3 + sumg * dexp3;
        // ---- compact asymmetric (CASYM95) ----
        const T u1 = T(CAM_SP(0.2210243144E+00)), u2 = T(CAM_SP(0.4367417579E+00)),
                u3 = T(CAM_SP(0.6994985432E-02)), u4 = T(CAM_SP(0.1491096501E+01)),
                u5 = T(CAM_SP(0.1602896673E+01)), u6 = T(CAM_SP(-.2821747323E+01)),
                u7 = T(CAM_SP(0.4948310833E+00)), u8 = T(CAM_SP(-.3540394679E-01)),
                u9 = T(CAM_SP(-.3305809954E+01)), u10 = T(CAM_SP(0.3644382172E+01)),
                u11 = T(CAM_SP(-.9997570970E+00)), u12 = T(CAM_SP(0.7989919534E-01)),
                u13 = T(CAM_SP(-.1075807322E-02));
        const T sumt = t[0] + t[1] + t[2];
        const T pr = r[0] * r[1] * r[2];
        const T ipr = T(1) / pr;
        const T ipr2 = ipr * ipr, ipr3 = ipr2 * ipr;
        const T a2 = a * a;
        const T s2 = u9 * ipr2 + u10 * ipr + u11 + u12 * pr + u13 * pr * pr;
        const T series = T(1) + u4 * ipr2 + u5 * ipr + u6 + u7 * pr + u8 * pr * pr + a * s2;
        const T term1 = u1 * rt::exp(-u2 * rt::log(pr));   // u1 / pr^u2
        const T eterm = rt::exp(-u3 * rsq * rs);
        const T base = term1 * series * eterm;
        v += sumt * a2 * base;
        const T dpr_all = -T(2) * u4 * ipr3 - u5 * ipr2 + u7 + T(2) * u8 * pr;
        const T dpr_a = -T(2) * u9 * ipr3 - u10 * ipr2 + u12 + T(2) * u13 * pr;
        const T dpr[3] = { r[1] * r[2], r[2] * r[0], r[0] * r[1] };
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const T dterm1 = -u2 * term1 * ipr * dpr[i];
            const T determ = eterm * (-T(3) * u3 * rsq);
            const T dseries = dpr[i] * dpr_all + da[i] * s2 + a * dpr[i] * dpr_a;
            dv[i] += dt[i] * a2 * base + T(2) * a * da[i] * sumt * base
                     + sumt * a2 * eterm * (dterm1 * series + dseries * term1)
                     + determ * sumt * a2 * term1 * series;
        }
    }
    {
        const double tb[3] = { double(t[0]), double(t[1]), double(t[2]) };
        const double dtb[3] = { double(dt[0]), double(dt[1]), double(dt[2]) };
        if (F64) {
            vd = double(v);
            dvd[0] = double(dv[0]);
            dvd[1] = double(dv[1]);
            dvd[2] = double(dv[2]);
        } else {
            vd += double(v);
            dvd[0] += double(dv[0]);
            dvd[1] += double(dv[1]);
            dvd[2] += double(dv[2]);
        }
        bend<double>(rd, icompc, tb, dtb, vd, dvd);
    }
}

// One bead: x = 9 coordinates in bohr (atom-major), returns V (Eh) and writes
// g = dV/dx (Eh/bohr).  Distances are formed in double from the double
// positions and then evaluated in T; the chain rule back to Cartesians and
// the outputs are double.  R(1)=|r_B-r_A|, R(2)=|r_A-r_C|, R(3)=|r_C-r_B|
// (src/egrad_h3.f:44-74).
template <typename T>
__device__ __forceinline__ void egrad_bead(const double x[9], double& v, double g[9])
{
    double ab[3], ac[3], bc[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        ab[d] = x[3 + d] - x[d];
        ac[d] = x[d] - x[6 + d];
        bc[d] = x[6 + d] - x[3 + d];
    }
    const double rab = ::sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]);
    const double rac = ::sqrt(ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2]);
    const double rbc = ::sqrt(bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2]);
    const double r[3] = { rab, rac, rbc };
    double dr[3];
    pote<T>(r, v, dr);
    const double f0 = dr[0] / rab, f1 = dr[1] / rac, f2 = dr[2] / rbc;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        g[d] = f1 * ac[d] - f0 * ab[d];
        g[3 + d] = f0 * ab[d] - f2 * bc[d];
        g[6 + d] = f2 * bc[d] - f1 * ac[d];
    }
}

#undef CAM_SP
}  // namespace h3
}  // namespace cam
