/*
 * oracle/qmdff.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Plain-C restatement of the reference's QMDFF energy and gradient as
 * gradient calls it for one QMDFF (src/gradient.f90:341-348): ff_eg
 * (src/ff_eg.f90:36-629: bond stretch, bend with abdamp, torsions and
 * out-of-plane terms), ff_nonb without Ewald (src/ff_nonb.f90:34-520: the
 * nci pair list, the intermolecular pairs of a multi-molecule box, Coulomb
 * plain, switched or Zahn) and ff_hb's hydrogen bonds (src/ff_hb.f90:36-280,
 * src/eabhag.f90).  Parameters are the arrays prepare leaves in the qmdff
 * module (1-based atom indices as the file gives them).  Loops and the order
 * of every sum follow the reference, and so do its quirks: the bend's single
 * precision abdamp constant, the unsigned torsion angle, Coulomb gradients of
 * q q / r^2 form also for the Zahn energy, eabhag's minimum image on the
 * A-H vector only.  Halogen bonds (eabxag) are not restated: a parameter set
 * with them is rejected (returns NAN).
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

#define SP(x) ((double)(float)(x))
static const double PI = 3.1415926535897932384626433832795029;
static const double PI2 = 6.28318530717958623199592693708837;
static const double SPI = 1.77245385090551599275151910313925;

/* box_image (src/box_image.f90:36-64): subtract box lengths while outside */
static void image(const oracle_qmdff *f, double *d)
{
    if (!f->periodic) return;
    for (int c = 0; c < 3; ++c) {
        const double L = f->box[c], L2 = 0.5 * f->box[c];
        while (fabs(d[c]) > L2) d[c] = d[c] - copysign(L, d[c]);
    }
}

static void sub(const double *x, int a, int b, double *d)
{
    for (int c = 0; c < 3; ++c) d[c] = x[3 * (a - 1) + c] - x[3 * (b - 1) + c];
}

static void cross(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static double vnorm(double *r, int normalise)     /* vecnorm_loc (src/vecnorm.f90) */
{
    double sp = 0.0;
    for (int i = 0; i < 3; ++i) sp = sp + r[i] * r[i];
    const double rn = sqrt(sp);
    if (normalise && fabs(rn) > 1e-14) {
        const double o = 1.0 / rn;
        for (int i = 0; i < 3; ++i) r[i] = o * r[i];
    }
    return rn;
}

static double vlen(const double *a)                /* src/vlen.f90 */
{
    const double t = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    return t > 0.0 ? sqrt(t) : 0.0;
}

/* abdamp (src/abdamp.f90:36-51): 3.0 * 3.5710642 is a default-REAL product */
static void abdamp(const oracle_qmdff *f, int ati, int atj, double r2, double *damp, double *ddamp)
{
    const double c = (double)(3.0f * 3.5710642f);
    const double s = f->rad[ati - 1] + f->rad[atj - 1];
    const double rcut = c * (s * s);
    const double rr = (r2 / rcut) * (r2 / rcut);
    *damp = 1.0 / (1.0 + rr);
    *ddamp = -2.0 * 2 * rr / (r2 * ((1.0 + rr) * (1.0 + rr)));
}

/* valijkl (src/valijkl.f90): the unsigned dihedral acos(n_a . n_b) */
static double valijkl(const oracle_qmdff *f, const double *x, int i, int j, int k, int l)
{
    double ra[3], rb[3], rc[3], na[3], nb[3];
    sub(x, j, i, ra);
    sub(x, k, j, rb);
    sub(x, l, k, rc);
    image(f, ra);
    image(f, rb);
    image(f, rc);
    cross(ra, rb, na);
    cross(rb, rc, nb);
    vnorm(na, 1);
    vnorm(nb, 1);
    double s = 0.0;
    for (int c = 0; c < 3; ++c) s = s + na[c] * nb[c];
    if (fabs(fabs(s) - 1.0) < 1e-14) s = copysign(1.0, s);
    return acos(s);
}

/* dphidr (src/dphidr.f90) */
static void dphidr(const oracle_qmdff *f, const double *x, int i, int j, int k, int l, double phi, double *di,
                   double *dj, double *dk, double *dl)
{
    double ra[3], rb[3], rc[3], rapb[3], rbpc[3], na[3], nb[3];
    double rab[3], rba[3], rac[3], rbb[3], rbc[3], raa[3], rapba[3], rapbb[3], rbpca[3], rbpcb[3];
    const double cosphi = cos(phi), sinphi = sin(phi);
    sub(x, j, i, ra);
    sub(x, k, j, rb);
    sub(x, l, k, rc);
    image(f, ra);
    image(f, rb);
    image(f, rc);
    for (int c = 0; c < 3; ++c) {
        rapb[c] = ra[c] + rb[c];
        rbpc[c] = rb[c] + rc[c];
    }
    cross(ra, rb, na);
    cross(rb, rc, nb);
    const double nan_ = vnorm(na, 0), nbn = vnorm(nb, 0);
    const double nenner = nan_ * nbn * sinphi;
    if (fabs(nenner) < 1e-14) {
        for (int c = 0; c < 3; ++c) di[c] = dj[c] = dk[c] = dl[c] = 0.0;
        return;
    }
    const double on = 1.0 / nenner;
    cross(na, rb, rab);
    cross(nb, ra, rba);
    cross(na, rc, rac);
    cross(nb, rb, rbb);
    cross(nb, rc, rbc);
    cross(na, ra, raa);
    cross(rapb, na, rapba);
    cross(rapb, nb, rapbb);
    cross(rbpc, na, rbpca);
    cross(rbpc, nb, rbpcb);
    for (int c = 0; c < 3; ++c) {
        di[c] = on * (cosphi * nbn / nan_ * rab[c] - rbb[c]);
        dj[c] = on * (cosphi * (nbn / nan_ * rapba[c] + nan_ / nbn * rbc[c]) - (rac[c] + rapbb[c]));
        dk[c] = on * (cosphi * (nbn / nan_ * raa[c] + nan_ / nbn * rbpcb[c]) - (rba[c] + rbpca[c]));
        dl[c] = on * (cosphi * nan_ / nbn * rbb[c] - rab[c]);
    }
}

/* omega (src/omega.f90): out-of-plane angle of l over the plane i, j, k */
static double omega(const oracle_qmdff *f, const double *x, int i, int j, int k, int l)
{
    double re[3], rd[3], rv[3], rn[3];
    sub(x, i, j, re);
    sub(x, k, j, rd);
    sub(x, l, i, rv);
    image(f, re);
    image(f, rd);
    image(f, rv);
    cross(re, rd, rn);
    vnorm(rn, 1);
    vnorm(rv, 1);
    const double rnv = rn[0] * rv[0] + rn[1] * rv[1] + rn[2] * rv[2];
    return asin(rnv);
}

/* domegadr (src/domegadr.f90) */
static void domegadr(const oracle_qmdff *f, const double *x, int i, int j, int k, int l, double om, double *di,
                     double *dj, double *dk, double *dl)
{
    double rv[3], rd[3], re[3], rdme[3], rn[3], rve[3], rne[3], rdv[3], rdn[3], rvdme[3], rndme[3];
    const double sinom = sin(om);
    sub(x, l, i, rv);
    sub(x, k, j, rd);
    sub(x, i, j, re);
    image(f, rv);
    image(f, rd);
    image(f, re);
    for (int c = 0; c < 3; ++c) rdme[c] = rd[c] - re[c];
    cross(re, rd, rn);
    const double rvn = vnorm(rv, 0), rnn = vnorm(rn, 0);
    cross(rv, re, rve);
    cross(rn, re, rne);
    cross(rd, rv, rdv);
    cross(rd, rn, rdn);
    cross(rv, rdme, rvdme);
    cross(rn, rdme, rndme);
    const double nenner = rnn * rvn * cos(om);
    if (fabs(nenner) > 1e-14) {
        const double on = 1.0 / nenner;
        for (int c = 0; c < 3; ++c) {
            di[c] = on * (rdv[c] - rn[c] - sinom * (rvn / rnn * rdn[c] - rnn / rvn * rv[c]));
            dj[c] = on * (rvdme[c] - sinom * rvn / rnn * rndme[c]);
            dk[c] = on * (rve[c] - sinom * rvn / rnn * rne[c]);
            dl[c] = on * (rn[c] - sinom * rnn / rvn * rv[c]);
        }
    } else {
        for (int c = 0; c < 3; ++c) di[c] = dj[c] = dk[c] = dl[c] = 0.0;
    }
}

static void addg(double *g, int a, const double *v)
{
    for (int c = 0; c < 3; ++c) g[3 * (a - 1) + c] = g[3 * (a - 1) + c] + v[c];
}

/* ff_eg: e and g are cleared first, as the reference does. */
static double ff_eg(const oracle_qmdff *f, const double *x, double *g)
{
    double e = 0.0;
    memset(g, 0, sizeof(double) * 3 * (size_t)f->n);
    for (int m = 0; m < f->nbond; ++m) {                               /* src/ff_eg.f90:87-163 */
        const int i = f->bond[2 * m], j = f->bond[2 * m + 1];
        double rb[3], ga[3];
        sub(x, i, j, rb);
        image(f, rb);
        const double r2 = rb[0] * rb[0] + rb[1] * rb[1] + rb[2] * rb[2];
        const double r = sqrt(r2);
        const double rij = f->vbond[3 * m], kij = f->vbond[3 * m + 1], aai = f->vbond[3 * m + 2];
        const double aai2 = aai / 2;
        e = e + kij * (1.0 + pow(rij / r, aai) - 2.0 * pow(rij / r, aai2));
        const double fac = aai * kij * (-pow(rij / r, aai) + pow(rij / r, aai2)) / r2;
        for (int c = 0; c < 3; ++c) ga[c] = fac * rb[c];
        addg(g, i, ga);
        for (int c = 0; c < 3; ++c) ga[c] = -ga[c];
        addg(g, j, ga);
    }
    for (int m = 0; m < f->nangl; ++m) {                               /* :167-299 */
        const int j = f->angl[3 * m], i = f->angl[3 * m + 1], k = f->angl[3 * m + 2];
        const double c0 = f->vangl[2 * m], kijk = f->vangl[2 * m + 1];
        double vab[3], vcb[3], vp[3], deda[3], dedc[3], t1[3], t2[3], ga[3], gb[3], gc[3];
        sub(x, i, j, vab);
        sub(x, k, j, vcb);
        image(f, vab);
        image(f, vcb);
        const double rab2 = vab[0] * vab[0] + vab[1] * vab[1] + vab[2] * vab[2];
        const double rcb2 = vcb[0] * vcb[0] + vcb[1] * vcb[1] + vcb[2] * vcb[2];
        cross(vcb, vab, vp);
        const double rp = vlen(vp) + 1e-14;
        double cosa;
        {   /* impsc */
            double rimp = 0.0;
            for (int c = 0; c < 3; ++c) rimp = rimp + vab[c] * vcb[c];
            const double al = vlen(vab), bl = vlen(vcb);
            cosa = (al > 0.0 && bl > 0.0) ? rimp / (al * bl) : 0.0;
        }
        cosa = fmin(1.0, fmax(-1.0, cosa));
        const double theta = acos(cosa);
        double dij, d2ij, djk, d2jk;
        abdamp(f, f->at[i - 1], f->at[j - 1], rab2, &dij, &d2ij);
        abdamp(f, f->at[k - 1], f->at[j - 1], rcb2, &djk, &d2jk);
        const double damp = dij * djk;
        double ea, deddt;
        if (PI - c0 < 1e-6) {
            const double dt = theta - c0;
            ea = kijk * (dt * dt);
            deddt = 2.0 * kijk * dt;
        } else {
            ea = kijk * ((cosa - cos(c0)) * (cosa - cos(c0)));
            deddt = 2.0 * kijk * sin(theta) * (cos(c0) - cosa);
        }
        e = e + ea * damp;
        cross(vab, vp, deda);
        const double rm1 = -deddt / (rab2 * rp);
        for (int c = 0; c < 3; ++c) deda[c] = deda[c] * rm1;
        cross(vcb, vp, dedc);
        const double rm2 = deddt / (rcb2 * rp);
        for (int c = 0; c < 3; ++c) dedc[c] = dedc[c] * rm2;
        for (int c = 0; c < 3; ++c) {
            const double dedb = deda[c] + dedc[c];
            t1[c] = ea * d2ij * djk * vab[c];
            t2[c] = ea * d2jk * dij * vcb[c];
            ga[c] = deda[c] * damp + t1[c];
            gb[c] = -dedb * damp - t1[c] - t2[c];
            gc[c] = dedc[c] * damp + t2[c];
        }
        addg(g, i, ga);
        addg(g, j, gb);
        addg(g, k, gc);
    }
    for (int m = 0; m < f->ntors; ++m) {                               /* :303-619 */
        const int *t = f->tors + 6 * m;
        const double *vt = f->vtors + 14 * m;
        const int i = t[0], j = t[1], k = t[2], l = t[3], nt = t[4];
        const double phi0 = vt[0];
        double vab[3], vcb[3], vdc[3], da[3], db[3], dc[3], dd[3], t1[3], t2[3], t3[3], gv[3];
        double et = 0.0, dij = 0.0, damp;
        if (t[5] != 2) {
            sub(x, i, j, vab);
            sub(x, j, k, vcb);
            sub(x, k, l, vdc);
            image(f, vab);
            image(f, vcb);
            image(f, vdc);
            const double rij = vab[0] * vab[0] + vab[1] * vab[1] + vab[2] * vab[2];
            const double rjk = vcb[0] * vcb[0] + vcb[1] * vcb[1] + vcb[2] * vcb[2];
            const double rkl = vdc[0] * vdc[0] + vdc[1] * vdc[1] + vdc[2] * vdc[2];
            double dpij, d2ij, dpjk, d2jk, dpkl, d2kl;
            abdamp(f, f->at[i - 1], f->at[j - 1], rij, &dpij, &d2ij);
            abdamp(f, f->at[k - 1], f->at[j - 1], rjk, &dpjk, &d2jk);
            abdamp(f, f->at[k - 1], f->at[l - 1], rkl, &dpkl, &d2kl);
            damp = dpjk * dpij * dpkl;
            const double phi = valijkl(f, x, i, j, k, l);
            dphidr(f, x, i, j, k, l, phi, da, db, dc, dd);
            int mm = 2;
            for (int it = 0; it < nt; ++it) {
                const double rn = vt[mm];
                const double dphi1 = phi - phi0, dphi2 = phi + phi0 - PI2;
                const double c1 = rn * dphi1 + vt[mm + 1], c2 = rn * dphi2 + vt[mm + 1];
                const double x1cos = cos(c1), x2cos = cos(c2), x1sin = sin(c1), x2sin = sin(c2);
                const double phipi = phi - PI;
                const double ef = erf(phipi);
                const double e1 = vt[mm + 2] * (1.0 + x1cos), e2 = vt[mm + 2] * (1.0 + x2cos);
                et = et + 0.5 * (1.0 - ef) * e1 + (0.5 + 0.5 * ef) * e2;
                const double expo = exp(-(phipi * phipi)) / SPI;
                dij = dij - expo * e1 - 0.5 * (1.0 - ef) * vt[mm + 2] * x1sin * rn + expo * e2 -
                      (0.5 + 0.5 * ef) * vt[mm + 2] * x2sin * rn;
                mm = mm + 3;
            }
            et = et * vt[1];
            dij = dij * vt[1] * damp;
            for (int c = 0; c < 3; ++c) {
                t1[c] = et * d2ij * dpjk * dpkl * vab[c];
                t2[c] = et * d2jk * dpij * dpkl * vcb[c];
                t3[c] = et * d2kl * dpij * dpjk * vdc[c];
            }
            for (int c = 0; c < 3; ++c) gv[c] = dij * da[c] + t1[c];
            addg(g, i, gv);
            for (int c = 0; c < 3; ++c) gv[c] = dij * db[c] - t1[c] + t2[c];
            addg(g, j, gv);
            for (int c = 0; c < 3; ++c) gv[c] = dij * dc[c] + t3[c] - t2[c];
            addg(g, k, gv);
            for (int c = 0; c < 3; ++c) gv[c] = dij * dd[c] - t3[c];
            addg(g, l, gv);
            e = e + et * damp;
        } else {
            sub(x, j, i, vab);
            sub(x, j, k, vcb);
            sub(x, j, l, vdc);
            image(f, vab);
            image(f, vcb);
            image(f, vdc);
            const double rij = vab[0] * vab[0] + vab[1] * vab[1] + vab[2] * vab[2];
            const double rjk = vcb[0] * vcb[0] + vcb[1] * vcb[1] + vcb[2] * vcb[2];
            const double rjl = vdc[0] * vdc[0] + vdc[1] * vdc[1] + vdc[2] * vdc[2];
            double dpij, d2ij, dpjk, d2jk, dpjl, d2jl;
            abdamp(f, f->at[i - 1], f->at[j - 1], rij, &dpij, &d2ij);
            abdamp(f, f->at[k - 1], f->at[j - 1], rjk, &dpjk, &d2jk);
            abdamp(f, f->at[j - 1], f->at[l - 1], rjl, &dpjl, &d2jl);
            damp = dpjk * dpij * dpjl;
            const double phi = omega(f, x, i, j, k, l);
            domegadr(f, x, i, j, k, l, phi, da, db, dc, dd);
            const double rn = vt[2];
            if (rn > 1e-6) {
                const double c1 = (phi - phi0) + PI;
                et = (1.0 + cos(c1)) * vt[1];
                dij = -sin(c1) * vt[1] * damp;
            } else {
                et = vt[1] * ((cos(phi) - cos(phi0)) * (cos(phi) - cos(phi0)));
                dij = 2.0 * vt[1] * sin(phi) * (cos(phi0) - cos(phi)) * damp;
            }
            for (int c = 0; c < 3; ++c) {
                t1[c] = et * d2ij * dpjk * dpjl * vab[c];
                t2[c] = et * d2jk * dpij * dpjl * vcb[c];
                t3[c] = et * d2jl * dpij * dpjk * vdc[c];
            }
            for (int c = 0; c < 3; ++c) gv[c] = dij * da[c] - t1[c];
            addg(g, i, gv);
            for (int c = 0; c < 3; ++c) gv[c] = dij * db[c] + t1[c] + t2[c] + t3[c];
            addg(g, j, gv);
            for (int c = 0; c < 3; ++c) gv[c] = dij * dc[c] - t2[c];
            addg(g, k, gv);
            for (int c = 0; c < 3; ++c) gv[c] = dij * dd[c] - t3[c];
            addg(g, l, gv);
            e = e + et * damp;
        }
    }
    return e;
}

static int T(int z) { return z - 1; }

/* dispersion and repulsion of one pair, scaled by s (src/ff_nonb.f90:118-175) */
static void disp_rep(const oracle_qmdff *f, int i1, int i2, const double *vab, double r2, double r, double s,
                     double *e, double *ga)
{
    const double oner = 1.0 / r;
    const int z1 = T(f->at[i1 - 1]), z2 = T(f->at[i2 - 1]);
    const double R0 = f->r094[z1 * 94 + z2];
    const double c6 = f->c6xy[(size_t)(i2 - 1) * f->n + (i1 - 1)];
    const double r4 = r2 * r2, r6 = r4 * r2;
    const double r06 = pow(R0, 6);
    const double t6 = r6 + r06, t8 = r6 * r2 + r06 * R0 * R0;
    const double c6t6 = c6 / t6, c6t8 = c6 / t8;
    double t27 = f->sr42[z1 * 94 + z2] * c6t8;
    double e0 = c6t6 + t27;
    *e = *e - e0 * s;
    double drij = s * (c6t6 * 6.0 * r4 / t6 + 8.0 * t27 * r6 / t8);
    for (int c = 0; c < 3; ++c) ga[c] = vab[c] * drij;
    if (r < 25) {
        const double xz = f->zab[z1 * 94 + z2], alpha = f->r0ab[z1 * 94 + z2];
        t27 = xz * exp(-alpha * r);
        e0 = t27 * oner;
        *e = *e + e0 * s;
        drij = s * t27 * (alpha * r + 1.0) * oner / r2;
        for (int c = 0; c < 3; ++c) ga[c] = ga[c] - vab[c] * drij;
    }
}

static double exp_switch(double x0, double x1, double x)
{
    const double xv = (x - x0) / (x1 - x0);
    return exp(1.0) * exp(1.0 / (xv - 1.0));
}

/* Coulomb of one pair (src/ff_nonb.f90:400-509); returns 1 with e0 set when
 * the pair is inside the cutoff */
static int coulomb(const oracle_qmdff *f, int i1, int i2, const double *vab, double r2, double r, double scale,
                   double *e0)
{
    double sw = 1.0;
    if (f->periodic) {
        if (r > f->coul_cut) return 0;
        if (!f->zahn && r > f->cut_low) sw = exp_switch(f->cut_low, f->coul_cut, r);
    }
    if (r > f->coul_cut) return 0;
    (void)vab;
    (void)r2;
    const double oner = 1.0 / r;
    if (f->zahn)
        *e0 = f->q[i1 - 1] * f->q[i2 - 1] * ((erfc(f->zahn_a * r) * oner) - f->zahn_par * (r - f->coul_cut));
    else
        *e0 = f->q[i1 - 1] * f->q[i2 - 1] * oner * scale * sw;
    return 1;
}

static double ff_nonb(const oracle_qmdff *f, const double *x, double *g)
{
    double e = 0.0;
    const double vdw_cut = f->periodic ? f->vdw_cut : 10.0 / 0.52917721092;
    if (f->nnci <= 1 && f->nmols == 0) return e;
    double vab[3], ga[3];
    for (int k = 0; k < f->nnci; ++k) {                                /* :88-193 */
        const int i1 = f->nci[3 * k], i2 = f->nci[3 * k + 1], nk = f->nci[3 * k + 2];
        sub(x, i1, i2, vab);
        image(f, vab);
        const double r2 = vab[0] * vab[0] + vab[1] * vab[1] + vab[2] * vab[2];
        const double r = sqrt(r2);
        if (f->periodic && r > vdw_cut) continue;
        disp_rep(f, i1, i2, vab, r2, r, f->eps2[nk - 1], &e, ga);
        addg(g, i1, ga);
        for (int c = 0; c < 3; ++c) ga[c] = -ga[c];
        addg(g, i2, ga);
    }
    if (f->nmols > 1) {                                               /* :198-332 */
        for (int i = 1; i <= f->n - 1; ++i)
            for (int j = i + 1; j <= f->n; ++j) {
                if (f->molnum[i - 1] == f->molnum[j - 1]) continue;
                sub(x, i, j, vab);
                image(f, vab);
                const double r2 = vab[0] * vab[0] + vab[1] * vab[1] + vab[2] * vab[2];
                const double r = sqrt(r2);
                if (f->periodic && r > vdw_cut) continue;
                disp_rep(f, i, j, vab, r2, r, 1.0, &e, ga);
                addg(g, i, ga);
                for (int c = 0; c < 3; ++c) ga[c] = -ga[c];
                addg(g, j, ga);
            }
    }
    for (int k = 0; k < f->nnci; ++k) {                                /* Coulomb, :400-449 */
        const int i1 = f->nci[3 * k], i2 = f->nci[3 * k + 1], nk = f->nci[3 * k + 2];
        sub(x, i1, i2, vab);
        image(f, vab);
        const double r2 = vab[0] * vab[0] + vab[1] * vab[1] + vab[2] * vab[2];
        const double r = sqrt(r2);
        double e0;
        if (!coulomb(f, i1, i2, vab, r2, r, f->eps1[nk - 1], &e0)) continue;
        e = e + e0;
        const double drij = e0 / r2;
        for (int c = 0; c < 3; ++c) ga[c] = -vab[c] * drij;
        addg(g, i1, ga);
        for (int c = 0; c < 3; ++c) ga[c] = -ga[c];
        addg(g, i2, ga);
    }
    if (f->nmols > 1) {                                               /* :453-509 */
        for (int i = 1; i <= f->n - 1; ++i)
            for (int j = i + 1; j <= f->n; ++j) {
                if (f->molnum[i - 1] == f->molnum[j - 1]) continue;
                sub(x, i, j, vab);
                image(f, vab);
                const double r2 = vab[0] * vab[0] + vab[1] * vab[1] + vab[2] * vab[2];
                const double r = sqrt(r2);
                double e0;
                if (!coulomb(f, i, j, vab, r2, r, 1.0, &e0)) continue;
                e = e + e0;
                const double drij = e0 / r2;
                for (int c = 0; c < 3; ++c) ga[c] = -vab[c] * drij;
                addg(g, i, ga);
                for (int c = 0; c < 3; ++c) ga[c] = -ga[c];
                addg(g, j, ga);
            }
    }
    return e;
}

/* eabhag (src/eabhag.f90): A...H-B with the minimum image on A-H only */
static double eabhag(const oracle_qmdff *f, const double *x, int A, int B, int H, double ca, double cb, double *g)
{
    const double longcut = 8.0, alp7 = 12.0, alp3 = 6.0;
    double drah[3], drbh[3], drab[3], ga[3], gb[3], gh[3], dg[3], dg2[3];
    sub(x, A, H, drah);
    sub(x, B, H, drbh);
    sub(x, A, B, drab);
    image(f, drah);
    const double rab2 = drab[0] * drab[0] + drab[1] * drab[1] + drab[2] * drab[2];
    const double rab = sqrt(rab2);
    const double rah2 = drah[0] * drah[0] + drah[1] * drah[1] + drah[2] * drah[2];
    const double rah = sqrt(rah2);
    const double rbh2 = drbh[0] * drbh[0] + drbh[1] * drbh[1] + drbh[2] * drbh[2];
    const double rbh = sqrt(rbh2);
    const double ratio = pow(rab / longcut, alp7);
    double rdampl = 1.0 / (1.0 + ratio);
    rdampl = rdampl / rab2 / rab;
    double aprod, cosabh;
    if (rah2 > rbh2) {
        aprod = 1.0 / rbh / rab;
        cosabh = -(drbh[0] * drab[0] + drbh[1] * drab[1] + drbh[2] * drab[2]) * aprod;
    } else {
        aprod = 1.0 / rah / rab;
        cosabh = (drah[0] * drab[0] + drah[1] * drab[1] + drah[2] * drab[2]) * aprod;
    }
    double aterm = 0.5 * (cosabh + 1.0);
    double apref = pow(aterm, alp3 - 1);
    aterm = aterm * apref;
    apref = alp3 * 0.5 * apref;
    const double rah4 = rah2 * rah2, rbh4 = rbh2 * rbh2;
    const double denom = 1.0 / (rah4 + rbh4);
    const double da = (ca * rah4 + cb * rbh4) * denom;
    const double eabh = -da * rdampl * aterm;
    if (eabh > -1e-8) return 0.0;
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
            dg[c] = -aprod * drbh[c] - cosabh * drab[c] / rab2;
            dg2[c] = aprod * drab[c] + cosabh * drbh[c] / rbh2;
            dg[c] = gi * dg[c];
            dg2[c] = gi * dg2[c];
            ga[c] = ga[c] + dg[c];
            gh[c] = gh[c] + dg2[c];
            gb[c] = gb[c] - dg[c] - dg2[c];
        }
    } else {
        for (int c = 0; c < 3; ++c) {
            dg[c] = -aprod * drah[c] + cosabh * drab[c] / rab2;
            dg2[c] = -aprod * drab[c] + cosabh * drah[c] / rah2;
            dg[c] = gi * dg[c];
            dg2[c] = gi * dg2[c];
            gb[c] = gb[c] + dg[c];
            gh[c] = gh[c] + dg2[c];
            ga[c] = ga[c] - dg[c] - dg2[c];
        }
    }
    addg(g, A, ga);
    addg(g, B, gb);
    addg(g, H, gh);
    return eabh;
}

static double hbpara(double a, double b, double q) { return exp(-a * q) / (exp(-a * q) + b); }

static int is_hb_acceptor(int z) { return z == 7 || z == 8 || z == 9 || z == 16 || z == 17; }
static int is_halogen(int z) { return z == 17 || z == 35 || z == 53 || z == 85; }

int oracle_qmdff_supported(const oracle_qmdff *f)
{
    for (int k = 0; k < f->nhb; ++k)
        if (f->at[f->hb[3 * k + 2] - 1] != 1) return 0;              /* X-bonds in the list */
    if (f->nmols > 1)
        for (int m = 0; m < f->nbond; ++m) {
            const int a = f->at[f->bond[2 * m] - 1], b = f->at[f->bond[2 * m + 1] - 1];
            if ((is_halogen(a) && b != 1) || (is_halogen(b) && a != 1)) return 0;
        }
    return 1;
}

static double ff_hb(const oracle_qmdff *f, const double *x, double *g, double e)
{
    double rb[3];
    if (f->nhb < 1 && f->nmols == 0) return e;
    for (int k = 0; k < f->nhb; ++k) {                                 /* src/ff_hb.f90:60-86 */
        const int i1 = f->hb[3 * k], i2 = f->hb[3 * k + 1];
        sub(x, i1, i2, rb);
        image(f, rb);
        const double r = sqrt(rb[0] * rb[0] + rb[1] * rb[1] + rb[2] * rb[2]);
        if (r > 15.0) continue;
        e = e + eabhag(f, x, i1, i2, f->hb[3 * k + 2], f->vhb[2 * k], f->vhb[2 * k + 1], g);
    }
    if (f->nmols > 1) {                                               /* :94-276, hydrogen bonds */
        const double rH = f->rad[0];
        for (int m = 0; m < f->nbond; ++m) {
            const int i1 = f->bond[2 * m], i2 = f->bond[2 * m + 1];
            int aH = 0, aA = 0;
            if (f->at[i1 - 1] == 1 && is_hb_acceptor(f->at[i2 - 1])) {
                aH = i1;
                aA = i2;
            }
            if (f->at[i2 - 1] == 1 && is_hb_acceptor(f->at[i1 - 1])) {
                aH = i2;
                aA = i1;
            }
            if (!aH) continue;
            for (int j = 1; j <= f->n; ++j) {
                if (f->molnum[aH - 1] == f->molnum[j - 1]) continue;
                const double cpar = f->scalehb[f->at[aA - 1] - 1] * f->scalehb[f->at[j - 1] - 1];
                if (!(cpar > 1e-6)) continue;
                sub(x, aA, aH, rb);
                image(f, rb);
                const double ri = sqrt(rb[0] * rb[0] + rb[1] * rb[1] + rb[2] * rb[2]);
                sub(x, j, aH, rb);
                image(f, rb);
                const double rj = sqrt(rb[0] * rb[0] + rb[1] * rb[1] + rb[2] * rb[2]);
                const double dum1 = SP(1.3) * (f->rad[f->at[aA - 1] - 1] + rH) / SP(0.52917726);
                const double dum2 = SP(1.3) * (f->rad[f->at[j - 1] - 1] + rH) / SP(0.52917726);
                if (!(ri < dum1 || rj < dum2)) continue;
                sub(x, aA, j, rb);
                image(f, rb);
                const double r = sqrt(rb[0] * rb[0] + rb[1] * rb[1] + rb[2] * rb[2]);
                if (r > 15.0) continue;
                const double c2 = hbpara(10.0, 5.0, f->q[aA - 1]) * f->scalehb[f->at[aA - 1] - 1];
                const double c1 = hbpara(10.0, 5.0, f->q[j - 1]) * f->scalehb[f->at[j - 1] - 1];
                e = e + eabhag(f, x, j, aA, aH, c1, c2, g);
            }
        }
    }
    return e;
}

/* One QMDFF structure: energy (hartree) and gradient g(3,n), as gradient
 * returns them for one QMDFF before its energy shift.  parts (may be NULL)
 * receives the running energy after ff_eg, after ff_nonb and after ff_hb.
 * NAN when the parameters hold halogen bonds, which are not restated. */
double oracle_qmdff_eval(const oracle_qmdff *f, const double *x, double *g, double *parts)
{
    if (!oracle_qmdff_supported(f)) return NAN;
    double e = ff_eg(f, x, g);
    if (parts) parts[0] = e;
    e = e + ff_nonb(f, x, g);
    if (parts) parts[1] = e;
    e = ff_hb(f, x, g, e);
    if (parts) parts[2] = e;
    return e;
}
