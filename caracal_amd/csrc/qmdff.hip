// The reference's QMDFF force field on the device (ff_eg + ff_nonb + ff_hb as
// src/gradient.f90:341-348 calls them for one QMDFF), for nbeads replicas of
// one structure at once: the box engine's force field for QMDFF boxes.
//
// Kernels of one evaluation (positions and gradients [bead][atom][3]):
//   qm_inter_kernel   pairs of atoms of different molecules (dispersion,
//                     repulsion, Coulomb; src/ff_nonb.f90:198-332, :453-509):
//                     one lane per atom i, j atoms tiled through LDS, the j
//                     range split over blocks; each lane sums its own i, so
//                     every pair is evaluated from both sides and no atomics
//                     are needed (partials [split][bead][atom][3]);
//   qm_hbacc_kernel,  intermolecular hydrogen bonds (src/ff_hb.f90:198-276):
//   qm_hbdon_kernel   every donor bond A-H against every acceptor atom j of
//                     another molecule, once from the acceptor's side (its
//                     gradient and the energy) and once from the donor's (the
//                     A and H gradients);
//   qm_gather_kernel  one lane per (bead, atom): the bonded terms, the nci
//                     pairs and the listed hydrogen bonds the atom takes part
//                     in (a term is evaluated once per atom it moves and its
//                     energy counted at its first atom), plus the partials
//                     above, in a fixed order;
//   qm_reduce_kernel  per-bead energies from the block partials in a fixed
//                     order.
// Results are bitwise reproducible run to run.  Arithmetic follows the
// reference routine by routine, quirks included (abdamp's single-precision
// constant, the unsigned torsion angle, q q / r^2 Coulomb gradients for the
// Zahn energy, eabhag's minimum image on the A-H vector only, hydrogen-bond
// thresholds in default-REAL literals).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "qmdff.hpp"

namespace cam {

constexpr int kQTile = 256;         // lanes per block; j atoms per LDS tile
constexpr int kQSplit = 8;          // blocks over the j range of one i tile
constexpr int kQMaxElem = 16;       // distinct elements of one force field
constexpr int kQImageCap = 64;      // minimum-image iterations (the reference loops until inside)

struct QDev {
    int n, ne, nbond, nangl, ntors, nnci, nhb, ndon, nacc;
    int periodic, zahn, inter;          // inter: more than one molecule
    double box[3], box2[3];
    double coul_cut, vdw_cut, cut_low, zahn_a, zahn_par;
    double eps1[6], eps2[6];
    const int *elem;                    // n: element slot 0..ne-1
    const int *mol;                     // n: molecule number
    const double *q;                    // n
    const double *c6;                   // n * n
    const double *r0, *r06, *sr42, *zab, *alpha;   // ne * ne: [slot_i][slot_j]
    const double *rad;                  // ne (bohr)
    const int *bond, *angl, *tors, *nci, *hb;      // 0-based atoms
    const double *vbond, *vangl, *vtors, *vhb;
    const int *rptr, *rcode;            // per atom: the term slots (and hydrogen-bond parts) it takes part in
    int sbase[5], nslot;                // first slot of bonds, bends, torsions, nci pairs, listed H-bonds
    const int *don, *acc;               // donor bonds (A, H) and acceptor atoms
    const double *don_lim, *acc_lim;    // hydrogen-bond distance thresholds
    const double *hbc;                  // n: hbpara(10, 5, q) * scalehb
    const double *scal;                 // n: scalehb of the atom's element
};

// term codes of the per-atom lists: kind << 28 | role << 26 | index
enum { kTBond = 0, kTAngle = 1, kTTors = 2, kTNci = 3, kTHbList = 4, kTHbAcc = 5, kTHbDon = 6 };

struct QmdffDev {
    QDev d{};
    std::vector<void *> allocs;
    double *gpart = nullptr, *epart = nullptr, *gacc = nullptr, *eacc = nullptr, *gdon = nullptr;
    double *gslot = nullptr, *eslot = nullptr;
    int cap_beads = 0;
};

namespace {

__device__ inline void qimg(const QDev &f, double *d)
{
    if (!f.periodic) return;
#pragma unroll
    for (int c = 0; c < 3; ++c)
        for (int it = 0; it < kQImageCap && fabs(d[c]) > f.box2[c]; ++it) d[c] = d[c] - copysign(f.box[c], d[c]);
}

__device__ inline void qsub(const double *x, int a, int b, double *d)
{
    d[0] = x[3 * a] - x[3 * b];
    d[1] = x[3 * a + 1] - x[3 * b + 1];
    d[2] = x[3 * a + 2] - x[3 * b + 2];
}

__device__ inline void qcross(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ inline double qnorm(double *r, bool normalise)
{
    const double sp = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
    const double rn = sqrt(sp);
    if (normalise && fabs(rn) > 1e-14) {
        const double o = 1.0 / rn;
        r[0] = o * r[0];
        r[1] = o * r[1];
        r[2] = o * r[2];
    }
    return rn;
}

__device__ inline double qdot(const double *a, const double *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

// abdamp (src/abdamp.f90:36-51)
__device__ inline void qabdamp(const QDev &f, int ea, int eb, double r2, double &damp, double &ddamp)
{
    const double c = (double)(3.0f * 3.5710642f);
    const double s = f.rad[ea] + f.rad[eb];
    const double rcut = c * (s * s);
    const double rr = (r2 / rcut) * (r2 / rcut);
    damp = 1.0 / (1.0 + rr);
    ddamp = -2.0 * 2 * rr / (r2 * ((1.0 + rr) * (1.0 + rr)));
}

// bond stretch m, gradient on its atom of the given role (src/ff_eg.f90:87-163)
__device__ void qbond(const QDev &f, const double *x, int m, int role, double &e, double *gv)
{
    const int i = f.bond[2 * m], j = f.bond[2 * m + 1];
    double rb[3];
    qsub(x, i, j, rb);
    qimg(f, rb);
    const double r2 = (rb[0] * rb[0] + rb[1] * rb[1]) + rb[2] * rb[2];
    const double r = sqrt(r2);
    const double rij = f.vbond[3 * m], kij = f.vbond[3 * m + 1], aai = f.vbond[3 * m + 2];
    const double aai2 = aai / 2;
    const double p1 = pow(rij / r, aai), p2 = pow(rij / r, aai2);
    e = kij * (1.0 + p1 - 2.0 * p2);
    const double fac = aai * kij * (-p1 + p2) / r2;
    const double sgn = role == 0 ? 1.0 : -1.0;
    for (int c = 0; c < 3; ++c) gv[c] = sgn * (fac * rb[c]);
}

// bend m (src/ff_eg.f90:167-299): vertex angl[0]; roles 0 vertex, 1 and 2 the outer atoms
__device__ void qangle(const QDev &f, const double *x, int m, int role, double &e, double *gv)
{
    const int j = f.angl[3 * m], i = f.angl[3 * m + 1], k = f.angl[3 * m + 2];
    const double c0 = f.vangl[2 * m], kijk = f.vangl[2 * m + 1];
    double vab[3], vcb[3], vp[3], deda[3], dedc[3];
    qsub(x, i, j, vab);
    qsub(x, k, j, vcb);
    qimg(f, vab);
    qimg(f, vcb);
    const double rab2 = (vab[0] * vab[0] + vab[1] * vab[1]) + vab[2] * vab[2];
    const double rcb2 = (vcb[0] * vcb[0] + vcb[1] * vcb[1]) + vcb[2] * vcb[2];
    qcross(vcb, vab, vp);
    const double tp = (vp[0] * vp[0] + vp[1] * vp[1]) + vp[2] * vp[2];
    const double rp = (tp > 0.0 ? sqrt(tp) : 0.0) + 1e-14;
    const double rimp = qdot(vab, vcb);
    const double al = rab2 > 0.0 ? sqrt(rab2) : 0.0, bl = rcb2 > 0.0 ? sqrt(rcb2) : 0.0;
    double cosa = (al > 0.0 && bl > 0.0) ? rimp / (al * bl) : 0.0;
    cosa = fmin(1.0, fmax(-1.0, cosa));
    const double theta = acos(cosa);
    double dij, d2ij, djk, d2jk;
    qabdamp(f, f.elem[i], f.elem[j], rab2, dij, d2ij);
    qabdamp(f, f.elem[k], f.elem[j], rcb2, djk, d2jk);
    const double damp = dij * djk;
    double ea, deddt;
    if (3.1415926535897932384626433832795029 - c0 < 1e-6) {
        const double dt = theta - c0;
        ea = kijk * (dt * dt);
        deddt = 2.0 * kijk * dt;
    } else {
        const double dc = cosa - cos(c0);
        ea = kijk * (dc * dc);
        deddt = 2.0 * kijk * sin(theta) * (cos(c0) - cosa);
    }
    e = ea * damp;
    qcross(vab, vp, deda);
    const double rm1 = -deddt / (rab2 * rp);
    qcross(vcb, vp, dedc);
    const double rm2 = deddt / (rcb2 * rp);
    for (int c = 0; c < 3; ++c) {
        const double a = deda[c] * rm1, cc = dedc[c] * rm2;
        const double t1 = ea * d2ij * djk * vab[c], t2 = ea * d2jk * dij * vcb[c];
        if (role == 1) gv[c] = a * damp + t1;
        else if (role == 0) gv[c] = -(a + cc) * damp - t1 - t2;
        else gv[c] = cc * damp + t2;
    }
}

// torsion or out-of-plane term m (src/ff_eg.f90:303-619, src/valijkl.f90,
// src/dphidr.f90, src/omega.f90, src/domegadr.f90); roles 0..3 = i, j, k, l
__device__ void qtors(const QDev &f, const double *x, int m, int role, double &e, double *gv)
{
    const int *t = f.tors + 6 * m;
    const double *vt = f.vtors + 14 * m;
    const int i = t[0], j = t[1], k = t[2], l = t[3], nt = t[4];
    const double phi0 = vt[0];
    const double PI = 3.1415926535897932384626433832795029, PI2 = 6.28318530717958623199592693708837;
    const double SPI = 1.77245385090551599275151910313925;
    double vab[3], vcb[3], vdc[3], dd[3];
    double et = 0.0, dij = 0.0;
    if (t[5] != 2) {
        qsub(x, i, j, vab);
        qsub(x, j, k, vcb);
        qsub(x, k, l, vdc);
        qimg(f, vab);
        qimg(f, vcb);
        qimg(f, vdc);
        double dpij, d2ij, dpjk, d2jk, dpkl, d2kl;
        qabdamp(f, f.elem[i], f.elem[j], (vab[0] * vab[0] + vab[1] * vab[1]) + vab[2] * vab[2], dpij, d2ij);
        qabdamp(f, f.elem[k], f.elem[j], (vcb[0] * vcb[0] + vcb[1] * vcb[1]) + vcb[2] * vcb[2], dpjk, d2jk);
        qabdamp(f, f.elem[k], f.elem[l], (vdc[0] * vdc[0] + vdc[1] * vdc[1]) + vdc[2] * vdc[2], dpkl, d2kl);
        const double damp = dpjk * dpij * dpkl;
        // valijkl: the unsigned dihedral
        double ra[3], rb[3], rc[3], na[3], nb[3];
        qsub(x, j, i, ra);
        qsub(x, k, j, rb);
        qsub(x, l, k, rc);
        qimg(f, ra);
        qimg(f, rb);
        qimg(f, rc);
        qcross(ra, rb, na);
        qcross(rb, rc, nb);
        double nau[3] = { na[0], na[1], na[2] }, nbu[3] = { nb[0], nb[1], nb[2] };
        qnorm(nau, true);
        qnorm(nbu, true);
        double sab = qdot(nau, nbu);
        if (fabs(fabs(sab) - 1.0) < 1e-14) sab = copysign(1.0, sab);
        const double phi = acos(sab);
        // dphidr, only the row of this role
        const double cosphi = cos(phi), sinphi = sin(phi);
        const double nan_ = qnorm(na, false), nbn = qnorm(nb, false);
        const double nenner = nan_ * nbn * sinphi;
        if (fabs(nenner) < 1e-14) {
            dd[0] = dd[1] = dd[2] = 0.0;
        } else {
            const double on = 1.0 / nenner;
            double u[3], v[3], w[3], z[3];
            if (role == 0) {
                qcross(na, rb, u);
                qcross(nb, rb, v);
                for (int c = 0; c < 3; ++c) dd[c] = on * (cosphi * nbn / nan_ * u[c] - v[c]);
            } else if (role == 1) {
                double rapb[3] = { ra[0] + rb[0], ra[1] + rb[1], ra[2] + rb[2] };
                qcross(rapb, na, u);     // rapba
                qcross(nb, rc, v);       // rbc
                qcross(na, rc, w);       // rac
                qcross(rapb, nb, z);     // rapbb
                for (int c = 0; c < 3; ++c)
                    dd[c] = on * (cosphi * (nbn / nan_ * u[c] + nan_ / nbn * v[c]) - (w[c] + z[c]));
            } else if (role == 2) {
                double rbpc[3] = { rb[0] + rc[0], rb[1] + rc[1], rb[2] + rc[2] };
                qcross(na, ra, u);       // raa
                qcross(rbpc, nb, v);     // rbpcb
                qcross(nb, ra, w);       // rba
                qcross(rbpc, na, z);     // rbpca
                for (int c = 0; c < 3; ++c)
                    dd[c] = on * (cosphi * (nbn / nan_ * u[c] + nan_ / nbn * v[c]) - (w[c] + z[c]));
            } else {
                qcross(nb, rb, u);       // rbb
                qcross(na, rb, v);       // rab
                for (int c = 0; c < 3; ++c) dd[c] = on * (cosphi * nan_ / nbn * u[c] - v[c]);
            }
        }
        int mm = 2;
        for (int it = 0; it < nt; ++it) {
            const double rn = vt[mm];
            const double dphi1 = phi - phi0, dphi2 = phi + phi0 - PI2;
            const double c1 = rn * dphi1 + vt[mm + 1], c2 = rn * dphi2 + vt[mm + 1];
            const double x1cos = cos(c1), x2cos = cos(c2), x1sin = sin(c1), x2sin = sin(c2);
            const double phipi = phi - PI;
            const double ef = erf(phipi);
            const double ea1 = vt[mm + 2] * (1.0 + x1cos), ea2 = vt[mm + 2] * (1.0 + x2cos);
            et = et + 0.5 * (1.0 - ef) * ea1 + (0.5 + 0.5 * ef) * ea2;
            const double expo = exp(-(phipi * phipi)) / SPI;
            dij = dij - expo * ea1 - 0.5 * (1.0 - ef) * vt[mm + 2] * x1sin * rn + expo * ea2 -
                  (0.5 + 0.5 * ef) * vt[mm + 2] * x2sin * rn;
            mm = mm + 3;
        }
        et = et * vt[1];
        dij = dij * vt[1] * damp;
        double t1[3], t2[3], t3[3];
        for (int c = 0; c < 3; ++c) {
            t1[c] = et * d2ij * dpjk * dpkl * vab[c];
            t2[c] = et * d2jk * dpij * dpkl * vcb[c];
            t3[c] = et * d2kl * dpij * dpjk * vdc[c];
        }
        for (int c = 0; c < 3; ++c) {
            if (role == 0) gv[c] = dij * dd[c] + t1[c];
            else if (role == 1) gv[c] = dij * dd[c] - t1[c] + t2[c];
            else if (role == 2) gv[c] = dij * dd[c] + t3[c] - t2[c];
            else gv[c] = dij * dd[c] - t3[c];
        }
        e = et * damp;
    } else {
        qsub(x, j, i, vab);
        qsub(x, j, k, vcb);
        qsub(x, j, l, vdc);
        qimg(f, vab);
        qimg(f, vcb);
        qimg(f, vdc);
        double dpij, d2ij, dpjk, d2jk, dpjl, d2jl;
        qabdamp(f, f.elem[i], f.elem[j], (vab[0] * vab[0] + vab[1] * vab[1]) + vab[2] * vab[2], dpij, d2ij);
        qabdamp(f, f.elem[k], f.elem[j], (vcb[0] * vcb[0] + vcb[1] * vcb[1]) + vcb[2] * vcb[2], dpjk, d2jk);
        qabdamp(f, f.elem[j], f.elem[l], (vdc[0] * vdc[0] + vdc[1] * vdc[1]) + vdc[2] * vdc[2], dpjl, d2jl);
        const double damp = dpjk * dpij * dpjl;
        // omega and domegadr
        double re[3], rd[3], rv[3], rn[3];
        qsub(x, i, j, re);
        qsub(x, k, j, rd);
        qsub(x, l, i, rv);
        qimg(f, re);
        qimg(f, rd);
        qimg(f, rv);
        qcross(re, rd, rn);
        double rnu[3] = { rn[0], rn[1], rn[2] }, rvu[3] = { rv[0], rv[1], rv[2] };
        qnorm(rnu, true);
        qnorm(rvu, true);
        const double om = asin(qdot(rnu, rvu));
        const double sinom = sin(om);
        const double rvn = qnorm(rv, false), rnn = qnorm(rn, false);
        const double nenner = rnn * rvn * cos(om);
        if (fabs(nenner) > 1e-14) {
            const double on = 1.0 / nenner;
            double u[3], v[3];
            double rdme[3] = { rd[0] - re[0], rd[1] - re[1], rd[2] - re[2] };
            if (role == 0) {
                qcross(rd, rv, u);       // rdv
                qcross(rd, rn, v);       // rdn
                for (int c = 0; c < 3; ++c) dd[c] = on * (u[c] - rn[c] - sinom * (rvn / rnn * v[c] - rnn / rvn * rv[c]));
            } else if (role == 1) {
                qcross(rv, rdme, u);     // rvdme
                qcross(rn, rdme, v);     // rndme
                for (int c = 0; c < 3; ++c) dd[c] = on * (u[c] - sinom * rvn / rnn * v[c]);
            } else if (role == 2) {
                qcross(rv, re, u);       // rve
                qcross(rn, re, v);       // rne
                for (int c = 0; c < 3; ++c) dd[c] = on * (u[c] - sinom * rvn / rnn * v[c]);
            } else {
                for (int c = 0; c < 3; ++c) dd[c] = on * (rn[c] - sinom * rnn / rvn * rv[c]);
            }
        } else {
            dd[0] = dd[1] = dd[2] = 0.0;
        }
        const double rnt = vt[2];
        if (rnt > 1e-6) {
            const double c1 = (om - phi0) + PI;
            et = (1.0 + cos(c1)) * vt[1];
            dij = -sin(c1) * vt[1] * damp;
        } else {
            const double dc = cos(om) - cos(phi0);
            et = vt[1] * (dc * dc);
            dij = 2.0 * vt[1] * sin(om) * (cos(phi0) - cos(om)) * damp;
        }
        double t1[3], t2[3], t3[3];
        for (int c = 0; c < 3; ++c) {
            t1[c] = et * d2ij * dpjk * dpjl * vab[c];
            t2[c] = et * d2jk * dpij * dpjl * vcb[c];
            t3[c] = et * d2jl * dpij * dpjk * vdc[c];
        }
        for (int c = 0; c < 3; ++c) {
            if (role == 0) gv[c] = dij * dd[c] - t1[c];
            else if (role == 1) gv[c] = dij * dd[c] + t1[c] + t2[c] + t3[c];
            else if (role == 2) gv[c] = dij * dd[c] - t2[c];
            else gv[c] = dij * dd[c] - t3[c];
        }
        e = et * damp;
    }
}

// Dispersion and repulsion of the pair (a, b) scaled by s, gradient on a for
// d = x_a - x_b (src/ff_nonb.f90:118-175, :228-320); T is the arithmetic.
template <typename T>
__device__ inline void qdisp_rep(const QDev &f, int ea, int eb, double c6d, const T *d, T r2, T r, T s, T &e, T *ga)
{
    const int pe = ea * f.ne + eb;
    const T R0 = (T)f.r0[pe];
    const T c6 = (T)c6d;
    const T r4 = r2 * r2, r6 = r4 * r2;
    const T r06 = (T)f.r06[pe];
    const T t6 = r6 + r06, t8 = r6 * r2 + r06 * R0 * R0;
    const T c6t6 = c6 / t6, c6t8 = c6 / t8;
    T t27 = (T)f.sr42[pe] * c6t8;
    e = -(c6t6 + t27) * s;
    T drij = s * (c6t6 * (T)6.0 * r4 / t6 + (T)8.0 * t27 * r6 / t8);
    for (int c = 0; c < 3; ++c) ga[c] = d[c] * drij;
    if (r < (T)25) {
        const T xz = (T)f.zab[pe], alpha = (T)f.alpha[pe];
        t27 = xz * exp(-alpha * r);
        const T oner = (T)1.0 / r;
        e = e + t27 * oner * s;
        drij = s * t27 * (alpha * r + (T)1.0) * oner / r2;
        for (int c = 0; c < 3; ++c) ga[c] = ga[c] - d[c] * drij;
    }
}

// Coulomb of the pair (src/ff_nonb.f90:400-509); false outside the cutoff
template <typename T>
__device__ inline bool qcoulomb(const QDev &f, double qa, double qb, const T *d, T r2, T r, T scale, T &e, T *ga)
{
    T sw = (T)1.0;
    if (f.periodic) {
        if (r > (T)f.coul_cut) return false;
        if (!f.zahn && r > (T)f.cut_low) {
            const T xv = (r - (T)f.cut_low) / ((T)f.coul_cut - (T)f.cut_low);
            sw = (T)exp(1.0) * exp((T)1.0 / (xv - (T)1.0));
        }
    }
    if (r > (T)f.coul_cut) return false;
    const T oner = (T)1.0 / r;
    if (f.zahn)
        e = (T)qa * (T)qb * ((erfc((T)f.zahn_a * r) * oner) - (T)f.zahn_par * (r - (T)f.coul_cut));
    else
        e = (T)qa * (T)qb * oner * scale * sw;
    const T drij = e / r2;
    for (int c = 0; c < 3; ++c) ga[c] = -d[c] * drij;
    return true;
}

// nci pair m, gradient on its atom of the given role (src/ff_nonb.f90:88-193, :400-449)
__device__ void qnci(const QDev &f, const double *x, int m, int role, double &e, double *gv)
{
    const int i1 = f.nci[3 * m], i2 = f.nci[3 * m + 1], nk = f.nci[3 * m + 2] - 1;
    double d[3];
    qsub(x, i1, i2, d);
    qimg(f, d);
    const double r2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
    const double r = sqrt(r2);
    const double vdw = f.periodic ? f.vdw_cut : 10.0 / 0.52917721092;
    double ga[3] = { 0.0, 0.0, 0.0 }, gc[3];
    e = 0.0;
    if (!(f.periodic && r > vdw)) qdisp_rep<double>(f, f.elem[i1], f.elem[i2], f.c6[(size_t)i2 * f.n + i1], d, r2, r,
                                                    f.eps2[nk], e, ga);
    double ec;
    if (qcoulomb<double>(f, f.q[i1], f.q[i2], d, r2, r, f.eps1[nk], ec, gc)) {
        e = e + ec;
        for (int c = 0; c < 3; ++c) ga[c] = ga[c] + gc[c];
    }
    const double sgn = role == 0 ? 1.0 : -1.0;
    for (int c = 0; c < 3; ++c) gv[c] = sgn * ga[c];
}

// eabhag (src/eabhag.f90): energy and the gradient of all three atoms
__device__ double qeabhag(const QDev &f, const double *x, int A, int B, int H, double ca, double cb, double *ga,
                          double *gb, double *gh)
{
    const double longcut = 8.0, alp7 = 12.0, alp3 = 6.0;
    double drah[3], drbh[3], drab[3], dg[3], dg2[3];
    qsub(x, A, H, drah);
    qsub(x, B, H, drbh);
    qsub(x, A, B, drab);
    qimg(f, drah);
    const double rab2 = (drab[0] * drab[0] + drab[1] * drab[1]) + drab[2] * drab[2];
    const double rab = sqrt(rab2);
    const double rah2 = (drah[0] * drah[0] + drah[1] * drah[1]) + drah[2] * drah[2];
    const double rah = sqrt(rah2);
    const double rbh2 = (drbh[0] * drbh[0] + drbh[1] * drbh[1]) + drbh[2] * drbh[2];
    const double rbh = sqrt(rbh2);
    const double ratio = pow(rab / longcut, alp7);
    double rdampl = 1.0 / (1.0 + ratio);
    rdampl = rdampl / rab2 / rab;
    double aprod, cosabh;
    if (rah2 > rbh2) {
        aprod = 1.0 / rbh / rab;
        cosabh = -qdot(drbh, drab) * aprod;
    } else {
        aprod = 1.0 / rah / rab;
        cosabh = qdot(drah, drab) * aprod;
    }
    double aterm = 0.5 * (cosabh + 1.0);
    double apref = pow(aterm, alp3 - 1);
    aterm = aterm * apref;
    apref = alp3 * 0.5 * apref;
    const double rah4 = rah2 * rah2, rbh4 = rbh2 * rbh2;
    const double denom = 1.0 / (rah4 + rbh4);
    const double da = (ca * rah4 + cb * rbh4) * denom;
    const double eabh = -da * rdampl * aterm;
    if (eabh > -1e-8) {
        for (int c = 0; c < 3; ++c) ga[c] = gb[c] = gh[c] = 0.0;
        return 0.0;
    }
    double gi = 4.0 * (ca - cb) * rah2 * rbh4 * denom * denom;
    gi = -gi * rdampl * aterm;
    for (int c = 0; c < 3; ++c) ga[c] = gi * drah[c];
    gi = 4.0 * (cb - ca) * rbh2 * rah4 * denom * denom;
    gi = -gi * rdampl * aterm;
    for (int c = 0; c < 3; ++c) gb[c] = gi * drbh[c];
    for (int c = 0; c < 3; ++c) gh[c] = -ga[c] - gb[c];
    gi = rdampl * rdampl * rab * (3.0 + (3.0 + alp7) * ratio);
    gi = gi * da * aterm;
    for (int c = 0; c < 3; ++c) {
        dg[c] = gi * drab[c];
        ga[c] = ga[c] + dg[c];
        gb[c] = gb[c] - dg[c];
    }
    gi = -da * rdampl * apref;
    if (rah2 > rbh2) {
        for (int c = 0; c < 3; ++c) {
            dg[c] = gi * (-aprod * drbh[c] - cosabh * drab[c] / rab2);
            dg2[c] = gi * (aprod * drab[c] + cosabh * drbh[c] / rbh2);
            ga[c] = ga[c] + dg[c];
            gh[c] = gh[c] + dg2[c];
            gb[c] = gb[c] - dg[c] - dg2[c];
        }
    } else {
        for (int c = 0; c < 3; ++c) {
            dg[c] = gi * (-aprod * drah[c] + cosabh * drab[c] / rab2);
            dg2[c] = gi * (-aprod * drab[c] + cosabh * drah[c] / rah2);
            gb[c] = gb[c] + dg[c];
            gh[c] = gh[c] + dg2[c];
            ga[c] = ga[c] - dg[c] - dg2[c];
        }
    }
    return eabh;
}

// Listed hydrogen bond m (src/ff_hb.f90:60-86); roles 0 = A, 1 = B, 2 = H
__device__ void qhblist(const QDev &f, const double *x, int m, int role, double &e, double *gv)
{
    const int a = f.hb[3 * m], b = f.hb[3 * m + 1], h = f.hb[3 * m + 2];
    double d[3], ga[3], gb[3], gh[3];
    qsub(x, a, b, d);
    qimg(f, d);
    const double r = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    e = 0.0;
    gv[0] = gv[1] = gv[2] = 0.0;
    if (r > 15.0) return;
    e = qeabhag(f, x, a, b, h, f.vhb[2 * m], f.vhb[2 * m + 1], ga, gb, gh);
    const double *s = role == 0 ? ga : role == 1 ? gb : gh;
    for (int c = 0; c < 3; ++c) gv[c] = s[c];
}

// Intermolecular hydrogen bond of donor bond dn against acceptor slot ac
// (src/ff_hb.f90:198-276): eabhag(j, A, H) when its distance conditions hold
__device__ double qhbinter(const QDev &f, const double *x, int dn, int ac, double *gj, double *ga, double *gh)
{
    const int A = f.don[2 * dn], H = f.don[2 * dn + 1], j = f.acc[ac];
    gj[0] = gj[1] = gj[2] = ga[0] = ga[1] = ga[2] = gh[0] = gh[1] = gh[2] = 0.0;
    if (f.mol[H] == f.mol[j]) return 0.0;
    const double cpar = f.scal[A] * f.scal[j];
    if (!(cpar > 1e-6)) return 0.0;
    double d[3];
    qsub(x, A, H, d);
    qimg(f, d);
    const double ri = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    qsub(x, j, H, d);
    qimg(f, d);
    const double rj = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    if (!(ri < f.don_lim[dn] || rj < f.acc_lim[ac])) return 0.0;
    qsub(x, A, j, d);
    qimg(f, d);
    const double r = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    if (r > 15.0) return 0.0;
    return qeabhag(f, x, j, A, H, f.hbc[j], f.hbc[A], gj, ga, gh);
}

// ---------------------------------------------------------------------------

template <typename T>
struct QAcc {
    double s = 0.0;
    __device__ void add(T v) { s = s + (double)v; }
    __device__ double get() const { return s; }
};
template <>
struct QAcc<float> {                    // compensated FP64 accumulation of binary32 terms
    double s = 0.0, c = 0.0;
    __device__ void add(float v)
    {
        const double y = (double)v - c;
        const double t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ double get() const { return s; }
};

template <typename T>
__global__ __launch_bounds__(kQTile) void qm_inter_kernel(QDev f, const double *__restrict__ xall, double *gpart,
                                                          double *epart)
{
    __shared__ double sx[3 * kQTile];
    __shared__ double sq[kQTile];
    __shared__ int se[kQTile], sm[kQTile];
    __shared__ double sred[kQTile / 64];
    const int n = f.n, b = blockIdx.z, split = blockIdx.y, nsplit = gridDim.y;
    const double *x = xall + (size_t)b * n * 3;
    const int i = blockIdx.x * kQTile + threadIdx.x;
    const bool act = i < n;
    double xi[3] = { 0.0, 0.0, 0.0 }, qi = 0.0;
    int ei = 0, mi = -1;
    if (act) {
        xi[0] = x[3 * i];
        xi[1] = x[3 * i + 1];
        xi[2] = x[3 * i + 2];
        qi = f.q[i];
        ei = f.elem[i];
        mi = f.mol[i];
    }
    const int chunk = (n + nsplit - 1) / nsplit;
    const int j0 = split * chunk, j1 = min(n, j0 + chunk);
    const T vdw = f.periodic ? (T)f.vdw_cut : (T)(10.0 / 0.52917721092);
    QAcc<T> gx, gy, gz, ee;
    for (int t0 = j0; t0 < j1; t0 += kQTile) {
        __syncthreads();
        const int jj = t0 + threadIdx.x;
        if (jj < j1) {
            sx[3 * threadIdx.x] = x[3 * jj];
            sx[3 * threadIdx.x + 1] = x[3 * jj + 1];
            sx[3 * threadIdx.x + 2] = x[3 * jj + 2];
            sq[threadIdx.x] = f.q[jj];
            se[threadIdx.x] = f.elem[jj];
            sm[threadIdx.x] = f.mol[jj];
        }
        __syncthreads();
        const int nt = min(kQTile, j1 - t0);
        if (!act) continue;
        for (int u = 0; u < nt; ++u) {
            if (sm[u] == mi) continue;                      // same molecule (and i itself)
            double dd[3] = { xi[0] - sx[3 * u], xi[1] - sx[3 * u + 1], xi[2] - sx[3 * u + 2] };
            qimg(f, dd);
            const T d[3] = { (T)dd[0], (T)dd[1], (T)dd[2] };
            const T r2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
            const T r = sqrt(r2);
            T ga[3] = { (T)0, (T)0, (T)0 }, e = (T)0;
            if (!(f.periodic && r > vdw))
                qdisp_rep<T>(f, ei, se[u], f.c6[(size_t)(t0 + u) * n + i], d, r2, r, (T)1.0, e, ga);
            T ec, gc[3];
            if (qcoulomb<T>(f, qi, sq[u], d, r2, r, (T)1.0, ec, gc)) {
                e = e + ec;
                ga[0] = ga[0] + gc[0];
                ga[1] = ga[1] + gc[1];
                ga[2] = ga[2] + gc[2];
            }
            gx.add(ga[0]);
            gy.add(ga[1]);
            gz.add(ga[2]);
            ee.add(e);
        }
    }
    if (act) {
        double *gp = gpart + (((size_t)split * gridDim.z + b) * n + i) * 3;
        gp[0] = gx.get();
        gp[1] = gy.get();
        gp[2] = gz.get();
    }
    // block energy (every pair counted from both sides: half of it)
    double v = act ? 0.5 * ee.get() : 0.0;
    for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kQTile / 64; ++w) s = s + sred[w];
        epart[((size_t)b * gridDim.x + blockIdx.x) * nsplit + split] = s;
    }
}

// Every (term, atom) slot of the bonded terms, nci pairs and listed hydrogen
// bonds: the term's gradient on that atom, and its energy at the slot of its
// first atom (block sums in a fixed order).  Slots are contiguous per kind.
__global__ __launch_bounds__(256) void qm_slot_kernel(QDev f, const double *__restrict__ xall, double *gslot,
                                                      double *eslot)
{
    __shared__ double sred[4];
    const int sl = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
    const double *x = xall + (size_t)b * f.n * 3;
    double e = 0.0;
    if (sl < f.nslot) {
        double et = 0.0, gv[3] = { 0.0, 0.0, 0.0 };
        int kind = 4;
        while (kind > 0 && sl < f.sbase[kind]) --kind;
        const int roles = kind == 0 ? 2 : kind == 1 ? 3 : kind == 2 ? 4 : kind == 3 ? 2 : 3;
        const int idx = (sl - f.sbase[kind]) / roles, role = (sl - f.sbase[kind]) % roles;
        switch (kind) {
        case kTBond: qbond(f, x, idx, role, et, gv); break;
        case kTAngle: qangle(f, x, idx, role, et, gv); break;
        case kTTors: qtors(f, x, idx, role, et, gv); break;
        case kTNci: qnci(f, x, idx, role, et, gv); break;
        default: qhblist(f, x, idx, role, et, gv); break;
        }
        double *o = gslot + ((size_t)b * f.nslot + sl) * 3;
        o[0] = gv[0];
        o[1] = gv[1];
        o[2] = gv[2];
        if (role == 0) e = et;
    }
    for (int o = 32; o > 0; o >>= 1) e = e + __shfl_xor(e, o);
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) eslot[(size_t)b * gridDim.x + blockIdx.x] = ((sred[0] + sred[1]) + sred[2]) + sred[3];
}

// acceptor side: one wave per acceptor atom, lanes over the donor bonds; the
// acceptor's gradient and the energy (butterfly sums: a fixed order)
__global__ __launch_bounds__(256) void qm_hbacc_kernel(QDev f, const double *__restrict__ xall, double *gacc,
                                                       double *eacc)
{
    const int lane = threadIdx.x & 63, ac = blockIdx.x * 4 + (threadIdx.x >> 6), b = blockIdx.y;
    if (ac >= f.nacc) return;
    const double *x = xall + (size_t)b * f.n * 3;
    double g[3] = { 0.0, 0.0, 0.0 }, e = 0.0, gj[3], ga[3], gh[3];
    for (int dn = lane; dn < f.ndon; dn += 64) {
        e = e + qhbinter(f, x, dn, ac, gj, ga, gh);
        g[0] = g[0] + gj[0];
        g[1] = g[1] + gj[1];
        g[2] = g[2] + gj[2];
    }
    for (int o = 32; o > 0; o >>= 1) {
        e = e + __shfl_xor(e, o);
        for (int c = 0; c < 3; ++c) g[c] = g[c] + __shfl_xor(g[c], o);
    }
    if (lane == 0) {
        double *out = gacc + ((size_t)b * f.nacc + ac) * 3;
        out[0] = g[0];
        out[1] = g[1];
        out[2] = g[2];
        eacc[(size_t)b * f.nacc + ac] = e;
    }
}

// donor side: one wave per donor bond, lanes over the acceptors; the A and H gradients
__global__ __launch_bounds__(256) void qm_hbdon_kernel(QDev f, const double *__restrict__ xall, double *gdon)
{
    const int lane = threadIdx.x & 63, dn = blockIdx.x * 4 + (threadIdx.x >> 6), b = blockIdx.y;
    if (dn >= f.ndon) return;
    const double *x = xall + (size_t)b * f.n * 3;
    double sa[3] = { 0.0, 0.0, 0.0 }, sh[3] = { 0.0, 0.0, 0.0 }, gj[3], ga[3], gh[3];
    for (int ac = lane; ac < f.nacc; ac += 64) {
        (void)qhbinter(f, x, dn, ac, gj, ga, gh);
        for (int c = 0; c < 3; ++c) {
            sa[c] = sa[c] + ga[c];
            sh[c] = sh[c] + gh[c];
        }
    }
    for (int o = 32; o > 0; o >>= 1)
        for (int c = 0; c < 3; ++c) {
            sa[c] = sa[c] + __shfl_xor(sa[c], o);
            sh[c] = sh[c] + __shfl_xor(sh[c], o);
        }
    if (lane == 0) {
        double *out = gdon + ((size_t)b * f.ndon + dn) * 6;
        for (int c = 0; c < 3; ++c) {
            out[c] = sa[c];
            out[3 + c] = sh[c];
        }
    }
}

// one lane per (bead, atom): its term slots, hydrogen-bond parts and
// intermolecular partials in a fixed order; the acceptors' energies
__global__ __launch_bounds__(256) void qm_gather_kernel(QDev f, int nsplit, int nbeads, const double *gslot,
                                                        const double *gpart, const double *gacc, const double *eacc,
                                                        const double *gdon, double *gout, double *egather)
{
    __shared__ double sred[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
    const bool act = i < f.n;
    double g[3] = { 0.0, 0.0, 0.0 }, e = 0.0;
    if (act) {
        for (int k = f.rptr[i]; k < f.rptr[i + 1]; ++k) {
            const int code = f.rcode[k];
            const int kind = code >> 28, payload = code & 0x0FFFFFFF;
            const double *src;
            if (kind == kTHbAcc) {
                src = gacc + ((size_t)b * f.nacc + payload) * 3;
                e = e + eacc[(size_t)b * f.nacc + payload];
            } else if (kind == kTHbDon) {
                const int role = payload >> 26, dn = payload & 0x03FFFFFF;
                src = gdon + ((size_t)b * f.ndon + dn) * 6 + (role == 1 ? 0 : 3);
            } else {
                src = gslot + ((size_t)b * f.nslot + payload) * 3;
            }
            g[0] = g[0] + src[0];
            g[1] = g[1] + src[1];
            g[2] = g[2] + src[2];
        }
        if (f.inter)
            for (int s = 0; s < nsplit; ++s) {
                const double *gp = gpart + (((size_t)s * nbeads + b) * f.n + i) * 3;
                g[0] = g[0] + gp[0];
                g[1] = g[1] + gp[1];
                g[2] = g[2] + gp[2];
            }
        if (gout) {
            double *o = gout + ((size_t)b * f.n + i) * 3;
            o[0] = g[0];
            o[1] = g[1];
            o[2] = g[2];
        }
    }
    double v = act ? e : 0.0;
    for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) egather[(size_t)b * gridDim.x + blockIdx.x] = ((sred[0] + sred[1]) + sred[2]) + sred[3];
}

__global__ void qm_reduce_kernel(const double *eslot, int nsl, const double *epart, int ninter, const double *egather,
                                 int ngather, double *e)
{
    const int b = blockIdx.x;
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int k = 0; k < nsl; ++k) s = s + eslot[(size_t)b * nsl + k];
    for (int k = 0; k < ngather; ++k) s = s + egather[(size_t)b * ngather + k];
    for (int k = 0; k < ninter; ++k) s = s + epart[(size_t)b * ninter + k];
    e[b] = s;
}

// src/verlet.f90:540-640 (orthorhombic box): an atom whose coordinate left
// [0, L] in some bead is moved with all its beads, bead by bead in order.
// One lane per (atom, coordinate): the beads are read in chunks of
// independent loads and the shifts recorded in order; the rare lane that
// shifts applies them to every bead in that order, as the reference adds them.
constexpr int kWrapChunk = 16, kWrapMaxShift = 8;

__global__ void qm_wrap_kernel(QDev f, double *x, int nbeads)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * f.n) return;
    const int i = t / 3, c = t % 3;
    const double L = f.box[c];
    // the shifts so far, in registers (every index below is a constant)
    double sh[kWrapMaxShift];
#pragma unroll
    for (int s = 0; s < kWrapMaxShift; ++s) sh[s] = 0.0;
    int ns = 0;
    bool over = false;
    auto push = [&](double d) {
        if (ns == kWrapMaxShift) {
            over = true;
            return;
        }
#pragma unroll
        for (int s = 0; s < kWrapMaxShift; ++s)
            if (s == ns) sh[s] = d;
        ++ns;
    };
    auto shifted = [&](double w) {
#pragma unroll
        for (int s = 0; s < kWrapMaxShift; ++s)
            if (s < ns) w = w + sh[s];
        return w;
    };
    for (int k0 = 0; k0 < nbeads && !over; k0 += kWrapChunk) {
        double v[kWrapChunk];
#pragma unroll
        for (int u = 0; u < kWrapChunk; ++u)
            v[u] = k0 + u < nbeads ? x[((size_t)(k0 + u) * f.n + i) * 3 + c] : 0.5 * L;
#pragma unroll
        for (int u = 0; u < kWrapChunk; ++u) {
            if (k0 + u >= nbeads || over) continue;
            double w = shifted(v[u]);
            while (w < 0.0 && !over) {
                push(L);
                if (!over) w = w + L;
            }
            while (w > L && !over) {
                push(-L);
                if (!over) w = w - L;
            }
        }
    }
    if (over) {                                 // many shifts: the reference's loop on memory
        for (int k = 0; k < nbeads; ++k) {
            int tries = 0;
            while (x[((size_t)k * f.n + i) * 3 + c] < 0.0 && tries <= 100) {
                for (int kk = 0; kk < nbeads; ++kk) x[((size_t)kk * f.n + i) * 3 + c] += L;
                ++tries;
            }
            while (x[((size_t)k * f.n + i) * 3 + c] > L && tries <= 100) {
                for (int kk = 0; kk < nbeads; ++kk) x[((size_t)kk * f.n + i) * 3 + c] -= L;
                ++tries;
            }
        }
        return;
    }
    if (ns == 0) return;
    for (int k = 0; k < nbeads; ++k) {
        double *p = x + ((size_t)k * f.n + i) * 3 + c;
        *p = shifted(*p);
    }
}

template <typename Tv>
int upload(QmdffDev *q, const std::vector<Tv> &v, const Tv **out)
{
    *out = nullptr;
    if (v.empty()) return CAM_OK;
    void *p = nullptr;
    CAM_HIP(hipMalloc(&p, v.size() * sizeof(Tv)));
    q->allocs.push_back(p);
    CAM_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(Tv), hipMemcpyHostToDevice));
    *out = static_cast<const Tv *>(p);
    return CAM_OK;
}

}  // namespace

void qmdff_free(QmdffDev *q)
{
    if (!q) return;
    for (void *p : q->allocs) (void)hipFree(p);
    double *bufs[] = { q->gpart, q->epart, q->gacc, q->eacc, q->gdon, q->gslot, q->eslot };
    for (double *p : bufs)
        if (p) (void)hipFree(p);
    delete q;
}

int qmdff_natoms(const QmdffDev *q) { return q->d.n; }

int qmdff_build(const cam_qmdff_desc *in, QmdffDev **out)
{
    *out = nullptr;
    const int n = in->natoms;
    std::unique_ptr<QmdffDev> q(new QmdffDev());
    QDev &d = q->d;
    d.n = n;
    d.nbond = in->nbond;
    d.nangl = in->nangl;
    d.ntors = in->ntors;
    d.nnci = in->nnci;
    d.nhb = in->nhb;
    d.periodic = in->periodic != 0;
    d.zahn = in->zahn != 0;
    for (int c = 0; c < 3; ++c) {
        d.box[c] = in->box[c];
        d.box2[c] = 0.5 * in->box[c];
    }
    d.coul_cut = in->coul_cut;
    d.vdw_cut = in->vdw_cut;
    d.cut_low = in->cut_low;
    d.zahn_a = in->zahn_a;
    d.zahn_par = in->zahn_par;
    for (int k = 0; k < 6; ++k) {
        d.eps1[k] = in->eps1[k];
        d.eps2[k] = in->eps2[k];
    }
    // element slots
    std::vector<int> zs;
    for (int i = 0; i < n; ++i)
        if (std::find(zs.begin(), zs.end(), in->at[i]) == zs.end()) zs.push_back(in->at[i]);
    CAM_CHECK_ARG((int)zs.size() <= kQMaxElem, "more than %d elements", kQMaxElem);
    const int ne = (int)zs.size();
    d.ne = ne;
    std::vector<int> elem(n), mol(n);
    std::vector<double> scal(n), hbc(n);
    for (int i = 0; i < n; ++i) {
        elem[i] = (int)(std::find(zs.begin(), zs.end(), in->at[i]) - zs.begin());
        mol[i] = (int)in->molnum[i];
        scal[i] = in->scalehb[in->at[i] - 1];
        const double eq = std::exp(-10.0 * in->q[i]);
        hbc[i] = eq / (eq + 5.0) * scal[i];       // hbpara(10, 5, q) * scalehb (src/ff_hb.f90:256-259)
    }
    std::vector<double> r0(ne * ne), r06(ne * ne), sr42(ne * ne), zab(ne * ne), alpha(ne * ne), rad(ne);
    for (int a = 0; a < ne; ++a) {
        rad[a] = in->rad[zs[a] - 1];
        for (int b = 0; b < ne; ++b) {
            const size_t fz = (size_t)(zs[b] - 1) * 94 + (zs[a] - 1);   // Fortran (z_a, z_b)
            r0[a * ne + b] = in->r094[fz];
            r06[a * ne + b] = std::pow(in->r094[fz], 6);
            sr42[a * ne + b] = in->sr42[fz];
            zab[a * ne + b] = in->zab[fz];
            alpha[a * ne + b] = in->r0ab[fz];
        }
    }
    // 0-based term lists
    auto zero_based = [](const int32_t *p, size_t cnt) {
        std::vector<int> v(cnt);
        for (size_t k = 0; k < cnt; ++k) v[k] = p[k] - 1;
        return v;
    };
    std::vector<int> bond = zero_based(in->bond, 2 * (size_t)in->nbond);
    std::vector<int> angl = zero_based(in->angl, 3 * (size_t)in->nangl);
    std::vector<int> tors(6 * (size_t)in->ntors), nci(3 * (size_t)in->nnci), hb = zero_based(in->hb, 3 * (size_t)in->nhb);
    for (int m = 0; m < in->ntors; ++m)
        for (int c = 0; c < 6; ++c) tors[6 * m + c] = c < 4 ? in->tors[6 * m + c] - 1 : in->tors[6 * m + c];
    for (int m = 0; m < in->nnci; ++m) {
        nci[3 * m] = in->nci[3 * m] - 1;
        nci[3 * m + 1] = in->nci[3 * m + 1] - 1;
        nci[3 * m + 2] = in->nci[3 * m + 2];
    }
    int nm = 0;
    bool multi = false;
    for (int i = 0; i < n; ++i) nm = std::max(nm, mol[i]);
    multi = nm > 1;                                   // nmols = maxval(molnum) (src/rdsolvff.f90:76)
    d.inter = multi;
    // intermolecular hydrogen bonds: donor bonds and acceptor atoms
    auto acceptor_z = [](int z) { return z == 7 || z == 8 || z == 9 || z == 16 || z == 17; };
    std::vector<int> don, acc;
    std::vector<double> don_lim, acc_lim;
    const float k13 = 1.3f, kbohr = 0.52917726f;
    if (multi) {
        for (int m = 0; m < in->nbond; ++m) {
            const int i1 = bond[2 * m], i2 = bond[2 * m + 1];
            int aH = -1, aA = -1;
            if (in->at[i1] == 1 && acceptor_z(in->at[i2])) {
                aH = i1;
                aA = i2;
            }
            if (in->at[i2] == 1 && acceptor_z(in->at[i1])) {
                aH = i2;
                aA = i1;
            }
            if (aH < 0) continue;
            don.push_back(aA);
            don.push_back(aH);
            don_lim.push_back((double)k13 * (in->rad[in->at[aA] - 1] + in->rad[0]) / (double)kbohr);
        }
        for (int j = 0; j < n; ++j)
            if (scal[j] > 0.0) {
                acc.push_back(j);
                acc_lim.push_back((double)k13 * (in->rad[in->at[j] - 1] + in->rad[0]) / (double)kbohr);
            }
    }
    if (don_lim.empty() || acc.empty()) {          // no pair to evaluate
        don.clear();
        don_lim.clear();
        acc.clear();
        acc_lim.clear();
    }
    d.ndon = (int)don_lim.size();
    d.nacc = (int)acc.size();
    // Term slots: one per (term, atom of the term), contiguous per kind;
    // each atom lists its slots (kind << 28 | slot) and its hydrogen-bond
    // parts (kind << 28 | role << 26 | index) in the reference's term order.
    const int64_t nslot64 = 2 * (int64_t)in->nbond + 3 * (int64_t)in->nangl + 4 * (int64_t)in->ntors +
                            2 * (int64_t)in->nnci + 3 * (int64_t)in->nhb;
    if (nslot64 >= (int64_t)1 << 28) return CAM_ERR_ARG;
    d.sbase[0] = 0;
    d.sbase[1] = 2 * in->nbond;
    d.sbase[2] = d.sbase[1] + 3 * in->nangl;
    d.sbase[3] = d.sbase[2] + 4 * in->ntors;
    d.sbase[4] = d.sbase[3] + 2 * in->nnci;
    d.nslot = (int)nslot64;
    std::vector<std::vector<int>> roles(n);
    auto code = [](int kind, int role, int idx) { return kind << 28 | role << 26 | idx; };
    auto slot = [&](int kind, int nroles, int m, int r) { return kind << 28 | (d.sbase[kind] + nroles * m + r); };
    for (int m = 0; m < in->nbond; ++m)
        for (int r = 0; r < 2; ++r) roles[bond[2 * m + r]].push_back(slot(kTBond, 2, m, r));
    for (int m = 0; m < in->nangl; ++m)
        for (int r = 0; r < 3; ++r) roles[angl[3 * m + r]].push_back(slot(kTAngle, 3, m, r));
    for (int m = 0; m < in->ntors; ++m)
        for (int r = 0; r < 4; ++r) roles[tors[6 * m + r]].push_back(slot(kTTors, 4, m, r));
    for (int m = 0; m < in->nnci; ++m)
        for (int r = 0; r < 2; ++r) roles[nci[3 * m + r]].push_back(slot(kTNci, 2, m, r));
    for (int m = 0; m < in->nhb; ++m)
        for (int r = 0; r < 3; ++r) roles[hb[3 * m + r]].push_back(slot(kTHbList, 3, m, r));
    for (int a = 0; a < d.nacc; ++a) roles[acc[a]].push_back(code(kTHbAcc, 0, a));
    for (int dn = 0; dn < d.ndon; ++dn) {
        roles[don[2 * dn]].push_back(code(kTHbDon, 1, dn));
        roles[don[2 * dn + 1]].push_back(code(kTHbDon, 2, dn));
    }
    std::vector<int> rptr(n + 1, 0), rcode;
    for (int i = 0; i < n; ++i) {
        rptr[i + 1] = rptr[i] + (int)roles[i].size();
        rcode.insert(rcode.end(), roles[i].begin(), roles[i].end());
    }
    std::vector<double> c6(in->c6xy, in->c6xy + (size_t)n * n);
    std::vector<double> qv(in->q, in->q + n);
    std::vector<double> vbond(in->vbond, in->vbond + 3 * (size_t)in->nbond);
    std::vector<double> vangl(in->vangl, in->vangl + 2 * (size_t)in->nangl);
    std::vector<double> vtors(in->vtors, in->vtors + 14 * (size_t)in->ntors);
    std::vector<double> vhb(in->vhb, in->vhb + 2 * (size_t)in->nhb);
    QmdffDev *qq = q.get();
    int rc = CAM_OK;
#define CAM_UP(vec, field)                                   \
    if (rc == CAM_OK) rc = upload(qq, vec, &d.field);
    CAM_UP(elem, elem) CAM_UP(mol, mol) CAM_UP(qv, q) CAM_UP(c6, c6) CAM_UP(r0, r0) CAM_UP(r06, r06)
    CAM_UP(sr42, sr42) CAM_UP(zab, zab) CAM_UP(alpha, alpha) CAM_UP(rad, rad) CAM_UP(bond, bond)
    CAM_UP(angl, angl) CAM_UP(tors, tors) CAM_UP(nci, nci) CAM_UP(hb, hb) CAM_UP(vbond, vbond)
    CAM_UP(vangl, vangl) CAM_UP(vtors, vtors) CAM_UP(vhb, vhb) CAM_UP(rptr, rptr) CAM_UP(rcode, rcode)
    CAM_UP(don, don) CAM_UP(acc, acc) CAM_UP(don_lim, don_lim) CAM_UP(acc_lim, acc_lim) CAM_UP(hbc, hbc)
    CAM_UP(scal, scal)
#undef CAM_UP
    if (rc != CAM_OK) return rc;
    *out = q.release();
    return CAM_OK;
}

namespace {
int qmdff_ensure(QmdffDev *q, int nbeads)
{
    if (nbeads <= q->cap_beads) return CAM_OK;
    double **bufs[] = { &q->gpart, &q->epart, &q->gacc, &q->eacc, &q->gdon, &q->gslot, &q->eslot };
    for (double **p : bufs)
        if (*p) {
            CAM_HIP(hipFree(*p));
            *p = nullptr;
        }
    const QDev &d = q->d;
    const int nbx = (d.n + kQTile - 1) / kQTile;
    const int ngx = (d.n + 255) / 256;
    CAM_HIP(hipMalloc(&q->gpart, (size_t)kQSplit * nbeads * d.n * 3 * sizeof(double)));
    CAM_HIP(hipMalloc(&q->epart, ((size_t)nbeads * nbx * kQSplit + (size_t)nbeads * ngx) * sizeof(double)));
    CAM_HIP(hipMalloc(&q->gacc, ((size_t)std::max(1, d.nacc) * nbeads * 4) * sizeof(double)));
    CAM_HIP(hipMalloc(&q->eacc, ((size_t)std::max(1, d.nacc) * nbeads) * sizeof(double)));
    CAM_HIP(hipMalloc(&q->gdon, ((size_t)std::max(1, d.ndon) * nbeads * 6) * sizeof(double)));
    CAM_HIP(hipMalloc(&q->gslot, ((size_t)std::max(1, d.nslot) * nbeads * 3) * sizeof(double)));
    CAM_HIP(hipMalloc(&q->eslot, ((size_t)std::max(1, (d.nslot + 255) / 256) * nbeads) * sizeof(double)));
    q->cap_beads = nbeads;
    return CAM_OK;
}
}  // namespace

template <typename T>
int qmdff_launch(QmdffDev *q, const double *x, int nbeads, double *e, double *g, hipStream_t s)
{
    int rc = qmdff_ensure(q, nbeads);
    if (rc != CAM_OK) return rc;
    const QDev &d = q->d;
    const int nbx = (d.n + kQTile - 1) / kQTile;
    const int ngx = (d.n + 255) / 256;
    double *egather = q->epart + (size_t)nbeads * nbx * kQSplit;
    if (d.inter) {
        hipLaunchKernelGGL(qm_inter_kernel<T>, dim3(nbx, kQSplit, nbeads), dim3(kQTile), 0, s, d, x, q->gpart,
                           q->epart);
        CAM_LAUNCH_CHECK();
    }
    if (d.nacc > 0 && d.ndon > 0) {
        hipLaunchKernelGGL(qm_hbacc_kernel, dim3((d.nacc + 3) / 4, nbeads), dim3(256), 0, s, d, x, q->gacc, q->eacc);
        CAM_LAUNCH_CHECK();
        hipLaunchKernelGGL(qm_hbdon_kernel, dim3((d.ndon + 3) / 4, nbeads), dim3(256), 0, s, d, x, q->gdon);
        CAM_LAUNCH_CHECK();
    }
    const int nsx = (d.nslot + 255) / 256;
    if (nsx > 0) {
        hipLaunchKernelGGL(qm_slot_kernel, dim3(nsx, nbeads), dim3(256), 0, s, d, x, q->gslot, q->eslot);
        CAM_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(qm_gather_kernel, dim3(ngx, nbeads), dim3(256), 0, s, d, kQSplit, nbeads, q->gslot, q->gpart,
                       q->gacc, q->eacc, q->gdon, g, egather);
    CAM_LAUNCH_CHECK();
    hipLaunchKernelGGL(qm_reduce_kernel, dim3(nbeads), dim3(64), 0, s, q->eslot, nsx, q->epart,
                       d.inter ? nbx * kQSplit : 0, egather, ngx, e);
    CAM_LAUNCH_CHECK();
    return CAM_OK;
}

template int qmdff_launch<double>(QmdffDev *, const double *, int, double *, double *, hipStream_t);
template int qmdff_launch<float>(QmdffDev *, const double *, int, double *, double *, hipStream_t);

int qmdff_wrap(QmdffDev *q, double *x, int nbeads, hipStream_t s)
{
    if (!q->d.periodic) return CAM_OK;
    const int nt = 3 * q->d.n;
    hipLaunchKernelGGL(qm_wrap_kernel, dim3((nt + 255) / 256), dim3(256), 0, s, q->d, x, nbeads);
    CAM_LAUNCH_CHECK();
    return CAM_OK;
}

}  // namespace cam
