// CH5 (CH4 + H) surface of Corchado, Bravo and Espinosa-Garcia on the device
// (reference: src/egrad_ch4h.f:33-1888 and the POTLIB glue src/util_ch4h.f).
//
// One lane evaluates one bead.  The reference keeps ~600 intermediates in
// COMMON blocks (theta0, dtheta0, dfdc, dfdh, ...); here everything that is a
// product of switching functions is recomputed where it is used so the bead
// lives in registers.  Index arithmetic is resolved at compile time by full
// unrolling; the one data-dependent choice of the reference -- swapping k and
// l in the out-of-plane term when the normal points the wrong way
// (src/egrad_ch4h.f:821-833) -- is applied as a sign on the normal, which is
// the same function: (a x b) is exactly antisymmetric under the swap.
//
// T is the arithmetic type (double, or float for the mixed path); positions
// and the gradient accumulators are always double, so FP32 terms are summed
// in FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include "quad.hpp"
#include "rt_math.hpp"

namespace cam {
namespace ch4h {

// Atom roles in the 6-atom input (src/egrad_ch4h.f:1852-1854): C is atom 2,
// the attacking H is atom 6, the methane H's h1..h4 are atoms 3,4,5,1.
constexpr int kC = 1, kB = 5;
__device__ __forceinline__ constexpr int hidx(int i) { return i == 3 ? 0 : i + 2; }

struct Par {
    // PREPOT's unit conversion applied (energies in 1e5 J/mol, lengths in angstrom).
    static constexpr double fact1 = 0.041840, fact2 = 6.022045;
    static constexpr double r0ch = 1.08898, a1ch = 1.78374, b1ch = 0.14201, c1ch = 2.21773;
    static constexpr double r0hh = 0.74239, ahh = 1.9589, r0cb = 1.08898, acb = 1.4173200;
    static constexpr double a3s = 0.1475300, b3s = -2.9926300;
    static constexpr double aphi = 0.5307000, bphi = 0.4012200, cphi = 1.9235100;
    static constexpr double atheta = 0.9119800, btheta = 0.3537500, ctheta = 1.8970500;
    static constexpr double aa1 = 1.265960, aa2 = 0.000710, aa3 = 0.985920, aa4 = 2.785060;
    static constexpr double a1s = 1.5132681e-7, b1s = -4.3792246, a2s = 1.9202402e-7, b2s = -12.323018;
};
__device__ __forceinline__ double d1ch() { return 111.266 * Par::fact1; }
__device__ __forceinline__ double d3ch() { return 48.96226 * Par::fact1; }
__device__ __forceinline__ double d1hh() { return 108.382 * Par::fact1; }
__device__ __forceinline__ double d3hh() { return 38.42657 * Par::fact1; }
__device__ __forceinline__ double d1cb() { return 56.505 * Par::fact1; }
__device__ __forceinline__ double d3cb() { return 19.612 * Par::fact1; }
__device__ __forceinline__ double fch3() { return 0.0693700 * Par::fact2; }
__device__ __forceinline__ double hch3() { return 0.1387400 * Par::fact2; }
__device__ __forceinline__ double fkinf() { return 0.4291400 * Par::fact2; }
__device__ __forceinline__ double akk() { return 0.1353000 * Par::fact2; }

// The two H's of {0..3} \ {i, j}, ascending.
__device__ __forceinline__ constexpr int comp0(int i, int j) { return (i != 0 && j != 0) ? 0 : (i != 1 && j != 1) ? 1 : 2; }
__device__ __forceinline__ constexpr int comp1(int i, int j) { return (i != 3 && j != 3) ? 3 : (i != 2 && j != 2) ? 2 : 1; }

// Value of lane r + n of the quad (frame index n); n is a compile-time
// constant wherever it is used, so this folds to one DPP move.
template <typename T>
__device__ __forceinline__ T fr(T v, int n)
{
    switch (n & 3) {
    case 0: return v;
    case 1: return quad::rot<1>(v);
    case 2: return quad::rot<2>(v);
    default: return quad::rot<3>(v);
    }
}

// theta0(i,j) and d theta0(i,j) / d rch(k) (refangles_ch4h, src/egrad_ch4h.f:367-501)
template <typename T, class V>
__device__ __forceinline__ T theta0(const V& s, int i, int j)
{
    const T tau = T(1.9106332362490186);           // acos(-1/3)
    const T a = tau - T(1.5707963267948966), b = tau - T(2.0943951023931953);
    const int k0 = comp0(i, j), l0 = comp1(i, j);
    return tau + a * (s.sphi(i) * s.sphi(j) - T(1)) + b * (s.sth(k0) * s.sth(l0) - T(1));
}
template <typename T, class V>
__device__ __forceinline__ T dtheta0(const V& s, int i, int j, int k)
{
    const T tau = T(1.9106332362490186);
    const T a = tau - T(1.5707963267948966), b = tau - T(2.0943951023931953);
    const int k0 = comp0(i, j), l0 = comp1(i, j);
    if (k == i) return a * s.dsphi(i) * s.sphi(j);
    if (k == j) return a * s.sphi(i) * s.dsphi(j);
    if (k == k0) return b * s.dsth(k0) * s.sth(l0);
    return b * s.sth(k0) * s.dsth(l0);
}

// Out-of-plane switch sw(i) = (1 - s3_i) prod_{m != i} s3_m and d sw(i)/d rch(j).
template <typename T, class V>
__device__ __forceinline__ T opswitch(const V& s, int i)
{
    const int a = (i + 1) & 3, b = (i + 2) & 3, c = (i + 3) & 3;
    return (T(1) - s.s3(i)) * s.s3(a) * s.s3(b) * s.s3(c);
}
template <typename T, class V>
__device__ __forceinline__ T dopswitch(const V& s, int i, int j)
{
    const int a = (i + 1) & 3, b = (i + 2) & 3, c = (i + 3) & 3;
    if (j == i) return -s.ds3(i) * s.s3(a) * s.s3(b) * s.s3(c);
    const T o = T(1) - s.s3(i);
    if (j == a) return o * s.ds3(a) * s.s3(b) * s.s3(c);
    if (j == b) return o * s.s3(a) * s.ds3(b) * s.s3(c);
    return o * s.s3(a) * s.s3(b) * s.ds3(c);
}

// In-plane force constant fk0(i,j) and its derivative wrt rch(k) (ipforce_ch4h).
template <typename T, class V>
__device__ __forceinline__ T fk0(const V& s, int i, int j)
{
    const T f0 = T(fkinf() + akk()), f2 = T(fkinf());
    const int k0 = comp0(i, j), l0 = comp1(i, j);
    return f0 + f0 * (s.s1(i) * s.s1(j) - T(1)) + (f0 - f2) * (s.s2(k0) * s.s2(l0) - T(1));
}
template <typename T, class V>
__device__ __forceinline__ T dfk0(const V& s, int i, int j, int k)
{
    const T f0 = T(fkinf() + akk()), f2 = T(fkinf());
    const int k0 = comp0(i, j), l0 = comp1(i, j);
    if (k == i) return f0 * s.ds1(i) * s.s1(j);
    if (k == j) return f0 * s.s1(i) * s.ds1(j);
    if (k == k0) return (f0 - f2) * s.ds2(k0) * s.s2(l0);
    return (f0 - f2) * s.s2(k0) * s.ds2(l0);
}

// ---------------------------------------------------------------------------
// Four lanes per bead.  Lane r (= lane id mod 4) owns methane H number r
// (atom hidx(r)); lane 0 also owns C and lane 1 owns the attacking H (Hb).
// Each lane evaluates, in its own rotated frame (frame index n is methane H
// (r + n) mod 4, fetched from lane r + n of the quad with DPP):
//   - the switching functions of its own H,
//   - the LEPS stretch of bond C-H_r / Hb-H_r,
//   - the out-of-plane term centred on H_r (opbend_ch4h rotation i = r),
//   - the in-plane pair (r, r+1), and half of the pair (r, r+2).
// Every factor that depends on a bond length only is accumulated as a scalar
// dV/dr; the angular terms are differentiated in vector form.  Results are
// reduced over the quad with DPP.  This is the reference's analytic gradient
// (src/egrad_ch4h.f:506-1540) rearranged; the rounding differs.
// ---------------------------------------------------------------------------
template <typename T>
struct Own {
    T t[3];            // t_ch = q_C - q_H (angstrom)
    T rch, rinv, rbh;
    T s1, ds1, s2, ds2, s3, ds3, sphi, dsphi, sth, dsth;
};

// 1 - tanh(x) and sech^2(x) from one exponential: with E = exp(2x),
// 1 - tanh(x) = 2/(1 + E) and sech^2(x) = (1 - tanh)(1 + tanh).  The reference
// evaluates tanh and cosh separately; this form needs one elementary function
// instead of two and has no cancellation where tanh is close to 1.
template <typename T>
__device__ __forceinline__ void one_minus_tanh(T x, T& s, T& sech2)
{
    const T e = rt::exp(T(2) * x);
    s = T(2) / (T(1) + e);
    sech2 = s * (T(2) - s);
}

template <typename T>
__device__ __forceinline__ void own_switches(Own<T>& o)
{
    const T r0ch = T(Par::r0ch);
    const T r = o.rch, dr = r - r0ch;
    {
        const T y = r - T(Par::b1s), y2 = y * y, y4 = y2 * y2, y8 = y4 * y4, y7 = y4 * y2 * y;
        const T arg = T(Par::a1s) * dr * y8;
        const bool on = arg < T(19);
        T sw, sh;
        one_minus_tanh(arg, sw, sh);
        o.s1 = on ? sw : T(0);
        o.ds1 = on ? -(T(Par::a1s) * (y8 + T(8) * dr * y7)) * sh : T(0);
    }
    {
        const T y = r - T(Par::b2s), y2 = y * y, y4 = y2 * y2, y6 = y4 * y2, y5 = y4 * y;
        const T arg = T(Par::a2s) * dr * y6;
        const bool on = arg < T(19);
        T sw, sh;
        one_minus_tanh(arg, sw, sh);
        o.s2 = on ? sw : T(0);
        o.ds2 = on ? -(T(Par::a2s) * (y6 + T(6) * dr * y5)) * sh : T(0);
    }
    {
        const T y = r - T(Par::b3s);
        const T arg = T(Par::a3s) * dr * (y * y);
        const bool on = arg < T(19);
        T sw, sh;
        one_minus_tanh(arg, sw, sh);
        o.s3 = on ? sw : T(0);
        o.ds3 = on ? -(T(Par::a3s) * (T(3) * r * r - T(2) * r * (r0ch + T(2) * T(Par::b3s))
                                      + T(Par::b3s) * (T(Par::b3s) + T(2) * r0ch))) * sh
                   : T(0);
    }
    const bool near = r < T(3.8);
    {
        const T y = r - T(Par::cphi);
        const T e = rt::exp(T(Par::bphi) * y * y * y);
        const T arg = T(Par::aphi) * dr * e;
        T sw, sh;
        one_minus_tanh(arg, sw, sh);
        o.sphi = near ? sw : T(0);
        o.dsphi = near ? -(T(Par::aphi) * (T(1) + T(3) * T(Par::bphi) * dr * y * y) * e) * sh : T(0);
    }
    {
        const T y = r - T(Par::ctheta);
        const T e = rt::exp(T(Par::btheta) * y * y * y);
        const T arg = T(Par::atheta) * dr * e;
        T sw, sh;
        one_minus_tanh(arg, sw, sh);
        o.sth = near ? sw : T(0);
        o.dsth = near ? -(T(Par::atheta) * (T(1) + T(3) * T(Par::btheta) * dr * y * y) * e) * sh : T(0);
    }
}

// The rotated frame: entry n is the quantity of methane H r + n, read from
// lane r + n of the quad with DPP where it is used rather than held in
// registers for the whole evaluation.
template <typename T>
struct Frame {
    const Own<T>& o;
    __device__ T t(int d, int n) const { return fr(o.t[d], n); }
    __device__ T rinv(int n) const { return fr(o.rinv, n); }
    __device__ T s1(int n) const { return fr(o.s1, n); }
    __device__ T ds1(int n) const { return fr(o.ds1, n); }
    __device__ T s2(int n) const { return fr(o.s2, n); }
    __device__ T ds2(int n) const { return fr(o.ds2, n); }
    __device__ T s3(int n) const { return fr(o.s3, n); }
    __device__ T ds3(int n) const { return fr(o.ds3, n); }
    __device__ T sphi(int n) const { return fr(o.sphi, n); }
    __device__ T dsphi(int n) const { return fr(o.dsphi, n); }
    __device__ T sth(int n) const { return fr(o.sth, n); }
    __device__ T dsth(int n) const { return fr(o.dsth, n); }
};

// Sum over the quad of rotated-frame contributions: lane m receives
// sum_o (value at frame index o on lane m - o), i.e. the total for its own H.
template <typename A>
__device__ __forceinline__ A unrotate_sum(const A (&loc)[4])
{
    return loc[0] + quad::rot<3>(loc[1]) + quad::rot<2>(loc[2]) + quad::rot<1>(loc[3]);
}

// In-plane pair (0, J) of the rotated frame (ipbend_ch4h / ipforce_ch4h),
// its energy and derivatives scaled by wt.
template <int J, typename T>
__device__ __forceinline__ void inplane_pair(const Frame<T>& R, const T (&f1)[4], const T (&df1dc)[4],
                                             const T (&df1dh)[4], T wt, double& vtot, T (&lg)[3][4], T (&ldc)[4],
                                             T (&ldh)[4])
{
    const T costh = (R.t(0, 0) * R.t(0, J) + R.t(1, 0) * R.t(1, J) + R.t(2, 0) * R.t(2, J)) * R.rinv(0) * R.rinv(J);
    const T dth = rt::acos(costh) - theta0<T>(R, 0, J);
    const T k0 = wt * fk0<T>(R, 0, J);
    const T kk = k0 * f1[0] * f1[J];
    vtot += double(T(0.5) * kk * dth * dth);
    const T pc = -kk * dth / rt::sqrt(T(1) - costh * costh);     // dV/d cos(theta)
    const T h2 = T(0.5) * dth * dth;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const T ui = -R.t(d, 0) * R.rinv(0), uj = -R.t(d, J) * R.rinv(J);
        lg[d][0] += pc * (uj - costh * ui) * R.rinv(0);
        lg[d][J] += pc * (ui - costh * uj) * R.rinv(J);
    }
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        T dk = wt * dfk0<T>(R, 0, J, n) * f1[0] * f1[J];
        if (n == 0) dk += k0 * df1dc[0] * f1[J];
        if (n == J) dk += k0 * f1[0] * df1dc[J];
        ldc[n] += -kk * dth * dtheta0<T>(R, 0, J, n) + h2 * dk;
    }
    ldh[0] += h2 * k0 * df1dh[0] * f1[J];
    ldh[J] += h2 * k0 * f1[0] * df1dh[J];
}

// One bead on a quad.  own[0..2]: position of methane H r (bohr); own[3..5]:
// position of C on lane 0, of Hb on lane 1 (ignored on lanes 2, 3).
// Returns the bead energy V (Eh, on all four lanes) and writes the gradient
// of the owned atoms (Eh/bohr) into gown (gown[3..5] is 0 on lanes 2, 3).
template <typename T>
__device__ __forceinline__ void egrad_quad(const double own[6], double& v_out, double gown[6])
{
    constexpr double toA = 0.52918;
    const int r = __lane_id() & 3;
    double qc[3], qb[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        qc[d] = quad::bcast<0>(own[3 + d]) * toA;
        qb[d] = quad::bcast<1>(own[3 + d]) * toA;
    }
    Own<T> o;
    T tcb[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        o.t[d] = T(qc[d] - own[d] * toA);
        tcb[d] = T(qc[d] - qb[d]);
    }
    const T rcb = rt::sqrt(tcb[0] * tcb[0] + tcb[1] * tcb[1] + tcb[2] * tcb[2]);
    o.rch = rt::sqrt(o.t[0] * o.t[0] + o.t[1] * o.t[1] + o.t[2] * o.t[2]);
    o.rinv = T(1) / o.rch;
    const T tb[3] = { o.t[0] - tcb[0], o.t[1] - tcb[1], o.t[2] - tcb[2] };   // q_Hb - q_H
    o.rbh = rt::sqrt(tb[0] * tb[0] + tb[1] * tb[1] + tb[2] * tb[2]);
    own_switches(o);

    double vtot = 0.0;
    T drch_own = T(0), drbh_own = T(0);     // direct scalar derivatives of the own bonds
    T lg[3][4] = {}, ldc[4] = {}, ldh[4] = {};

    // ---- LEPS stretch of the own bonds (stretch_ch4h, src/egrad_ch4h.f:506-708) ----
    T drcb, dach, dumach;
    {
        const T r0ch = T(Par::r0ch);
        const T rav = quad::sum(o.rch) * T(0.25);
        const T arga = T(Par::c1ch) * (rav - r0ch);
        T ach;
        if (arga < T(19)) {
            T sw, sh;
            one_minus_tanh(arga, sw, sh);
            ach = T(Par::a1ch) + T(Par::b1ch) * (T(2) - sw) * T(0.5);
            dumach = T(Par::b1ch) * T(Par::c1ch) * sh * T(0.5);
        } else {
            ach = T(Par::a1ch + Par::b1ch);
            dumach = T(0);
        }
        const T D1cb = T(d1cb()), D3cb = T(d3cb()), D1ch = T(d1ch()), D3ch = T(d3ch()),
                D1hh = T(d1hh()), D3hh = T(d3hh());
        const T xcb = rcb - T(Par::r0cb);
        const T ecb1 = rt::exp(-T(Par::acb) * xcb), ecb2 = ecb1 * ecb1;
        const T vqcb = T(0.5) * (D1cb * (ecb2 - T(2) * ecb1) + D3cb * (ecb2 + T(2) * ecb1));
        const T vjcb = T(0.5) * (D1cb * (ecb2 - T(2) * ecb1) - D3cb * (ecb2 + T(2) * ecb1));
        const T dqcb = -T(Par::acb) * ((D1cb + D3cb) * ecb2 - (D1cb - D3cb) * ecb1);
        const T djcb = -T(Par::acb) * ((D1cb - D3cb) * ecb2 - (D1cb + D3cb) * ecb1);
        const T xch = o.rch - r0ch, xbh = o.rbh - T(Par::r0hh);
        const T ech1 = rt::exp(-ach * xch), ech2 = ech1 * ech1;
        const T ebh1 = rt::exp(-T(Par::ahh) * xbh), ebh2 = ebh1 * ebh1;
        const T vqch = T(0.5) * (D1ch * (ech2 - T(2) * ech1) + D3ch * (ech2 + T(2) * ech1));
        const T vjch = T(0.5) * (D1ch * (ech2 - T(2) * ech1) - D3ch * (ech2 + T(2) * ech1));
        const T vqbh = T(0.5) * (D1hh * (ebh2 - T(2) * ebh1) + D3hh * (ebh2 + T(2) * ebh1));
        const T vjbh = T(0.5) * (D1hh * (ebh2 - T(2) * ebh1) - D3hh * (ebh2 + T(2) * ebh1));
        const T u = vjch - vjcb, w = vjcb - vjbh, z = vjbh - vjch;
        const T vj = -rt::sqrt((u * u + w * w + z * z) * T(0.5));
        vtot += double(vqch + vqcb + vqbh + vj);
        const T factj = T(0.5) / vj;
        const T ga = factj * (u - z), gbb = factj * (w - u), gc = factj * (z - w);
        const T gq = (D1ch + D3ch) * ech2 - (D1ch - D3ch) * ech1;
        const T gj = (D1ch - D3ch) * ech2 - (D1ch + D3ch) * ech1;
        drch_own = -ach * gq - ach * gj * ga;
        drbh_own = -T(Par::ahh) * ((D1hh + D3hh) * ebh2 - (D1hh - D3hh) * ebh1)
                   - T(Par::ahh) * ((D1hh - D3hh) * ebh2 - (D1hh + D3hh) * ebh1) * gc;
        dach = -xch * gq - xch * gj * ga;            // summed over the quad below
        drcb = dqcb + djcb * gbb;                     // summed over the quad below
    }

    const Frame<T> R{ o };

    // ---- out-of-plane bend centred on the own H (opbend_ch4h / calcdelta_ch4h) ----
    {
        const T Fch3 = T(fch3()), Hch3 = T(hch3());
        T a[3], b[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            a[d] = R.t(d, 1) - R.t(d, 2);         // q_Hk - q_Hj
            b[d] = R.t(d, 1) - R.t(d, 3);         // q_Hl - q_Hj
        }
        const T nx = a[1] * b[2] - a[2] * b[1], ny = a[2] * b[0] - a[0] * b[2], nz = a[0] * b[1] - a[1] * b[0];
        const T inrm = T(1) / rt::sqrt(nx * nx + ny * ny + nz * nz);
        T argd[3];
        int npos = 0;
#pragma unroll
        for (int ii = 0; ii < 3; ++ii) {
            const int m = ii + 1;
            argd[ii] = -(nx * R.t(0, m) + ny * R.t(1, m) + nz * R.t(2, m)) * R.rinv(m) * inrm;
            npos += argd[ii] > T(0) ? 1 : 0;
        }
        // the reference swaps k and l once per positive projection
        // (src/egrad_ch4h.f:821-833): the normal flips when the count is odd
        const T sg = (npos & 1) ? T(-1) : T(1);
        const T sw = opswitch<T>(R, 0);
        const T fi = Fch3 * sw, hi = Hch3 * sw;
        const T nh[3] = { sg * nx * inrm, sg * ny * inrm, sg * nz * inrm };
        T s2 = T(0), s4 = T(0);
        T vv[3] = { T(0), T(0), T(0) };
#pragma unroll
        for (int ii = 0; ii < 3; ++ii) {
            const int m = ii + 1;
            const T ad = sg * argd[ii];
            const T dm = rt::acos(ad) - theta0<T>(R, 0, m);
            const T d2 = dm * dm;
            s2 += d2;
            s4 += d2 * d2;
            const T wgt = T(2) * fi * dm + T(4) * hi * dm * d2;         // dV/d delta
#pragma unroll
            for (int n = 0; n < 4; ++n) ldc[n] -= wgt * dtheta0<T>(R, 0, m, n);
            const T pm = -wgt / rt::sqrt(T(1) - ad * ad);            // dV/d argd
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const T um = -R.t(d, m) * R.rinv(m);
                lg[d][m] += pm * (nh[d] - ad * um) * R.rinv(m);
                vv[d] += pm * sg * inrm * (um - ad * nh[d]);
            }
        }
        vtot += double(fi * s2 + hi * s4);
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            const T dsw = dopswitch<T>(R, 0, n);
            ldc[n] += (dsw * Fch3) * s2 + (dsw * Hch3) * s4;
        }
        const T bxv[3] = { b[1] * vv[2] - b[2] * vv[1], b[2] * vv[0] - b[0] * vv[2], b[0] * vv[1] - b[1] * vv[0] };
        const T vxa[3] = { vv[1] * a[2] - vv[2] * a[1], vv[2] * a[0] - vv[0] * a[2], vv[0] * a[1] - vv[1] * a[0] };
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lg[d][2] += bxv[d];
            lg[d][3] += vxa[d];
            lg[d][1] -= bxv[d] + vxa[d];
        }
    }

    // ---- in-plane pairs (r, r+1) and, on lanes 0 and 1, (r, r+2) ----
    {
        const T r0ch = T(Par::r0ch);
        T f1o, dco, dho;
        {
            const T rb = o.rbh, xch = o.rch - r0ch, xhh = rb - T(Par::r0hh);
            const T e1 = rt::exp(-T(Par::aa1) * rb * rb);
            const T e2 = rt::exp(-T(Par::aa4) * xhh * xhh);
            const T a1 = T(1) - e1;
            const T a2 = T(Par::aa2) + T(Par::aa3) * e2;
            const T eg = rt::exp(-a2 * (xch * xch));
            f1o = a1 * eg;
            dco = -T(2) * xch * a1 * a2 * eg;
            dho = T(2) * T(Par::aa1) * rb * e1 * eg + T(2) * T(Par::aa3) * T(Par::aa4) * xhh * e2 * (xch * xch) * a1 * eg;
        }
        T f1[4], df1dc[4], df1dh[4];
        f1[0] = f1o; df1dc[0] = dco; df1dh[0] = dho;
        f1[1] = quad::rot<1>(f1o); df1dc[1] = quad::rot<1>(dco); df1dh[1] = quad::rot<1>(dho);
        f1[2] = quad::rot<2>(f1o); df1dc[2] = quad::rot<2>(dco); df1dh[2] = quad::rot<2>(dho);
        f1[3] = quad::rot<3>(f1o); df1dc[3] = quad::rot<3>(dco); df1dh[3] = quad::rot<3>(dho);
        inplane_pair<1>(R, f1, df1dc, df1dh, T(1), vtot, lg, ldc, ldh);
        // the pair (r, r+2) is the pair (r+2, r) of lane r+2: every lane takes
        // half of it, so that the whole quad stays active for the DPP reads
        inplane_pair<2>(R, f1, df1dc, df1dh, T(0.5), vtot, lg, ldc, ldh);
    }

    // ---- quad reductions and the chain rule to Cartesians ----
    const T dachs = quad::sum(dach);
    const T drcbs = quad::sum(drcb);
    const T drch = drch_own + dachs * dumach * T(0.25) + unrotate_sum(ldc);
    const T drbh = drbh_own + unrotate_sum(ldh);
    T gh[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        T lgd[4] = { lg[d][0], lg[d][1], lg[d][2], lg[d][3] };
        gh[d] = unrotate_sum(lgd) - drch * o.t[d] * o.rinv - drbh * tb[d] / o.rbh;
    }
    const double scale = 0.0201723;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const T gbpart = quad::sum(drbh * tb[d] / o.rbh);
        const T gbd = gbpart - drcbs * tcb[d] / rcb;
        const T ghs = quad::sum(gh[d]);
        gown[d] = double(gh[d]) * scale;
        gown[3 + d] = r == 0 ? -double(ghs + gbd) * scale : r == 1 ? double(gbd) * scale : 0.0;
    }
    v_out = quad::sum(vtot) * 0.03812;
}

}  // namespace ch4h
}  // namespace cam
