/*
 * oracle/pes_ch4h.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * CPU restatement of the Corchado-Bravo-Espinosa-Garcia CH5 (CH4 + H)
 * valence-bond / molecular-mechanics surface as the reference evaluates it
 * (src/egrad_ch4h.f:33-1888 with the POTLIB glue in src/util_ch4h.f).
 * The reference keeps all intermediates in COMMON blocks; here they live in a
 * per-call struct so the routine is re-entrant, but the order of operations
 * follows the reference routine by routine.
 *
 * Atom order of the 6-atom input (src/egrad_ch4h.f:1852-1854): atom 2 is C,
 * atom 6 is the attacking H ("Hb"), atoms 3,4,5,1 are the methane H's h1..h4.
 * Units: q in bohr is scaled by 0.52918 to angstrom (:231), the energy in
 * 1e5 J/mol is scaled by 0.03812 to hartree (:267) and the gradient by
 * 0.0201723 to Eh/bohr (:276).  The POTLIB conversion factors (CNVRTD,
 * CNVRTE, CNVRTDE, ANUZERO) are all 1 / 0 for the flags the reference sets
 * (src/util_ch4h.f:680-719), so they are omitted.
 *
 * Parity: pinned against the compiled reference (oracle/_ref) and by central
 * differences of V (tests/test_oracle_pes.py).
 */
#include <math.h>
#include "oracle.h"

/* Parameters after PREPOT's unit conversion (src/egrad_ch4h.f:1781-1793,1855-1885). */
typedef struct {
    double r0ch, d1ch, d3ch, a1ch, b1ch, c1ch;
    double r0hh, d1hh, d3hh, ahh;
    double r0cb, d1cb, d3cb, acb;
    double a3s, b3s, aphi, bphi, cphi, atheta, btheta, ctheta;
    double fch3, hch3, fkinf, ak, bk, aa1, aa2, aa3, aa4;
} ch4h_par;

static ch4h_par make_par(void)
{
    const double fact1 = 0.041840, fact2 = 6.022045;
    ch4h_par p;
    p.r0ch = 1.08898;   p.d1ch = 111.266 * fact1;  p.d3ch = 48.96226 * fact1;
    p.a1ch = 1.78374;   p.b1ch = 0.14201;          p.c1ch = 2.21773;
    p.r0hh = 0.74239;   p.d1hh = 108.382 * fact1;  p.d3hh = 38.42657 * fact1;
    p.ahh = 1.9589;
    p.r0cb = 1.08898;   p.d1cb = 56.505 * fact1;   p.d3cb = 19.612 * fact1;
    p.acb = 1.4173200;
    p.a3s = 0.1475300;  p.b3s = -2.9926300;
    p.aphi = 0.5307000; p.bphi = 0.4012200;        p.cphi = 1.9235100;
    p.atheta = 0.9119800; p.btheta = 0.3537500;    p.ctheta = 1.8970500;
    p.fch3 = 0.0693700 * fact2;  p.hch3 = 0.1387400 * fact2;
    p.fkinf = 0.4291400 * fact2; p.ak = 0.1353000 * fact2;
    p.bk = 10.7132;
    p.aa1 = 1.265960; p.aa2 = 0.000710; p.aa3 = 0.985920; p.aa4 = 2.785060;
    return p;
}

/* x^n for small constant n by binary powering (how an integer power is expanded). */
static double ipow(double x, int n)
{
    double res = 0.0, sq = x;
    int have = 0;
    while (n) {
        if (n & 1) { res = have ? res * sq : sq; have = 1; }
        n >>= 1;
        if (n) sq = sq * sq;
    }
    return res;
}

typedef struct {
    double q[18], pdot[18];
    int nc[3], nhb[3], nh[4][3];                 /* 0-based coordinate indices */
    double tcb[3], tch[4][3], tbh[4][3], rcb, rch[4], rbh[4];
    double s1[4], ds1[4], s2[4], ds2[4], s3[4], ds3[4];
    double sphi[4], dsphi[4], stheta[4], dstheta[4];
    double theta0[4][4], dtheta0[4][4][4];
    double fdelta[4], hdelta[4], dfdelta[4][4], dhdelta[4][4];
    double fk0[4][4], f1[4], dfdc[4][4][4], dfdh[4][4][4];
} ch4h_ws;

static void set_index(ch4h_ws *w)
{
    /* nnc = 2, nnb = 6, nnh = (3,4,5,1): ncoord = 3*atom + ind - 3 (1-based). */
    const int nnh[4] = { 3, 4, 5, 1 };
    for (int ind = 0; ind < 3; ++ind) {
        w->nc[ind] = 3 * 2 - 3 + ind;
        w->nhb[ind] = 3 * 6 - 3 + ind;
        for (int i = 0; i < 4; ++i) w->nh[i][ind] = 3 * nnh[i] - 3 + ind;
    }
}

/* coorden_ch4h (src/egrad_ch4h.f:287-...) */
static void coorden(ch4h_ws *w)
{
    for (int ind = 0; ind < 3; ++ind) {
        w->tcb[ind] = w->q[w->nc[ind]] - w->q[w->nhb[ind]];
        for (int i = 0; i < 4; ++i) {
            w->tch[i][ind] = w->q[w->nc[ind]] - w->q[w->nh[i][ind]];
            w->tbh[i][ind] = w->q[w->nhb[ind]] - w->q[w->nh[i][ind]];
        }
    }
    w->rcb = sqrt(w->tcb[0] * w->tcb[0] + w->tcb[1] * w->tcb[1] + w->tcb[2] * w->tcb[2]);
    for (int i = 0; i < 4; ++i) {
        w->rch[i] = sqrt(w->tch[i][0] * w->tch[i][0] + w->tch[i][1] * w->tch[i][1] + w->tch[i][2] * w->tch[i][2]);
        w->rbh[i] = sqrt(w->tbh[i][0] * w->tbh[i][0] + w->tbh[i][1] * w->tbh[i][1] + w->tbh[i][2] * w->tbh[i][2]);
    }
}

/* switchf_ch4h (src/egrad_ch4h.f:1545-1685) */
static void switchf(ch4h_ws *w, const ch4h_par *p)
{
    const double a1s = 1.5132681e-7, b1s = -4.3792246, a2s = 1.9202402e-7, b2s = -12.323018;
    const double argmax = 19.0;
    for (int i = 0; i < 4; ++i) {
        const double r = w->rch[i];
        const double args1 = a1s * (r - p->r0ch) * ipow(r - b1s, 8);
        if (args1 < argmax) {
            w->s1[i] = 1.0 - tanh(args1);
            double d = a1s * (ipow(r - b1s, 8) + 8.0 * (r - p->r0ch) * ipow(r - b1s, 7));
            const double c = cosh(args1);
            w->ds1[i] = -d / (c * c);
        } else {
            w->s1[i] = 0.0;
            w->ds1[i] = 0.0;
        }
        const double args2 = a2s * (r - p->r0ch) * ipow(r - b2s, 6);
        if (args2 < argmax) {
            w->s2[i] = 1.0 - tanh(args2);
            double d = a2s * (ipow(r - b2s, 6) + 6.0 * (r - p->r0ch) * ipow(r - b2s, 5));
            const double c = cosh(args2);
            w->ds2[i] = -d / (c * c);
        } else {
            w->s2[i] = 0.0;
            w->ds2[i] = 0.0;
        }
        const double args3 = p->a3s * (r - p->r0ch) * ipow(r - p->b3s, 2);
        if (args3 < argmax) {
            w->s3[i] = 1.0 - tanh(args3);
            double d = p->a3s * (3.0 * ipow(r, 2) - 2.0 * r * (p->r0ch + 2.0 * p->b3s)
                                 + p->b3s * (p->b3s + 2.0 * p->r0ch));
            const double c = cosh(args3);
            w->ds3[i] = -d / (c * c);
        } else {
            w->s3[i] = 0.0;
            w->ds3[i] = 0.0;
        }
        if (r < 3.8) {
            const double e = exp(p->bphi * ipow(r - p->cphi, 3));
            const double arg = p->aphi * (r - p->r0ch) * e;
            w->sphi[i] = 1.0 - tanh(arg);
            double d = p->aphi * (1.0 + 3.0 * p->bphi * (r - p->r0ch) * ipow(r - p->cphi, 2));
            d = d * exp(p->bphi * ipow(r - p->cphi, 3));
            const double c = cosh(arg);
            w->dsphi[i] = -d / (c * c);
        } else {
            w->sphi[i] = 0.0;
            w->dsphi[i] = 0.0;
        }
        if (r < 3.8) {
            const double e = exp(p->btheta * ipow(r - p->ctheta, 3));
            const double arg = p->atheta * (r - p->r0ch) * e;
            w->stheta[i] = 1.0 - tanh(arg);
            double d = p->atheta * (1.0 + 3.0 * p->btheta * (r - p->r0ch) * ipow(r - p->ctheta, 2));
            d = d * exp(p->btheta * ipow(r - p->ctheta, 3));
            const double c = cosh(arg);
            w->dstheta[i] = -d / (c * c);
        } else {
            w->stheta[i] = 0.0;
            w->dstheta[i] = 0.0;
        }
    }
}

/* refangles_ch4h (src/egrad_ch4h.f:367-501) */
static void refangles(ch4h_ws *w)
{
    const double pi = 3.141592653589793;
    const double tau = acos(-1.0 / 3.0);
    const double halfpi = 0.5 * pi, twopi = 2.0 * pi;
    const double a = tau - halfpi, b = tau - twopi / 3.0;
    const double *sp = w->sphi, *dsp = w->dsphi, *st = w->stheta, *dst = w->dstheta;
    for (int i = 0; i < 4; ++i) {
        w->theta0[i][i] = 0.0;
        for (int k = 0; k < 4; ++k) w->dtheta0[i][i][k] = 0.0;
    }
    w->theta0[0][1] = tau + a * (sp[0] * sp[1] - 1.0) + b * (st[2] * st[3] - 1.0);
    w->theta0[0][2] = tau + a * (sp[0] * sp[2] - 1.0) + b * (st[1] * st[3] - 1.0);
    w->theta0[0][3] = tau + a * (sp[0] * sp[3] - 1.0) + b * (st[1] * st[2] - 1.0);
    w->theta0[1][2] = tau + a * (sp[1] * sp[2] - 1.0) + b * (st[0] * st[3] - 1.0);
    w->theta0[1][3] = tau + a * (sp[1] * sp[3] - 1.0) + b * (st[0] * st[2] - 1.0);
    w->theta0[2][3] = tau + a * (sp[2] * sp[3] - 1.0) + b * (st[0] * st[1] - 1.0);
    /* d/d rch(1) */
    w->dtheta0[0][1][0] = a * dsp[0] * sp[1];
    w->dtheta0[0][2][0] = a * dsp[0] * sp[2];
    w->dtheta0[0][3][0] = a * dsp[0] * sp[3];
    w->dtheta0[1][2][0] = b * dst[0] * st[3];
    w->dtheta0[1][3][0] = b * dst[0] * st[2];
    w->dtheta0[2][3][0] = b * dst[0] * st[1];
    /* d/d rch(2) */
    w->dtheta0[0][1][1] = a * sp[0] * dsp[1];
    w->dtheta0[0][2][1] = b * dst[1] * st[3];
    w->dtheta0[0][3][1] = b * dst[1] * st[2];
    w->dtheta0[1][2][1] = a * dsp[1] * sp[2];
    w->dtheta0[1][3][1] = a * dsp[1] * sp[3];
    w->dtheta0[2][3][1] = b * st[0] * dst[1];
    /* d/d rch(3) */
    w->dtheta0[0][1][2] = b * dst[2] * st[3];
    w->dtheta0[0][2][2] = a * sp[0] * dsp[2];
    w->dtheta0[0][3][2] = b * st[1] * dst[2];
    w->dtheta0[1][2][2] = a * sp[1] * dsp[2];
    w->dtheta0[1][3][2] = b * st[0] * dst[2];
    w->dtheta0[2][3][2] = a * dsp[2] * sp[3];
    /* d/d rch(4) */
    w->dtheta0[0][1][3] = b * st[2] * dst[3];
    w->dtheta0[0][2][3] = b * st[1] * dst[3];
    w->dtheta0[0][3][3] = a * sp[0] * dsp[3];
    w->dtheta0[1][2][3] = b * st[0] * dst[3];
    w->dtheta0[1][3][3] = a * sp[1] * dsp[3];
    w->dtheta0[2][3][3] = a * sp[2] * dsp[3];
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 4; ++j) {
            w->theta0[j][i] = w->theta0[i][j];
            for (int k = 0; k < 4; ++k) w->dtheta0[j][i][k] = w->dtheta0[i][j][k];
        }
}


/* stretch_ch4h: LEPS stretch with the tanh-switched C-H Morse exponent
 * (src/egrad_ch4h.f:506-708). */
static double stretch(ch4h_ws *w, const ch4h_par *p)
{
    double vqch[4], vjch[4], vqbh[4], vjbh[4], vq[4], vj[4], achdc[3], achdh[4][3];
    double *pd = w->pdot;
    const double rav = (w->rch[0] + w->rch[1] + w->rch[2] + w->rch[3]) / 4.0;
    double vstr = 0.0, ach, dumach;
    const double arga = p->c1ch * (rav - p->r0ch);
    if (arga < 19.0) {
        ach = p->a1ch + p->b1ch * (tanh(arga) + 1.0) * 0.5;
        const double c = cosh(arga);
        dumach = p->b1ch * p->c1ch / (2.0 * (c * c));
    } else {
        ach = p->a1ch + p->b1ch;
        dumach = 0.0;
    }
    const double xcb = w->rcb - p->r0cb;
    double e1 = p->d1cb * (exp(-2.0 * p->acb * xcb) - 2.0 * exp(-p->acb * xcb));
    double e3 = p->d3cb * (exp(-2.0 * p->acb * xcb) + 2.0 * exp(-p->acb * xcb));
    const double vqcb = (e1 + e3) * 0.5, vjcb = (e1 - e3) * 0.5;
    for (int i = 0; i < 4; ++i) {
        const double xch = w->rch[i] - p->r0ch, xbh = w->rbh[i] - p->r0hh;
        e1 = p->d1ch * (exp(-2.0 * ach * xch) - 2.0 * exp(-ach * xch));
        e3 = p->d3ch * (exp(-2.0 * ach * xch) + 2.0 * exp(-ach * xch));
        vqch[i] = (e1 + e3) * 0.5;
        vjch[i] = (e1 - e3) * 0.5;
        e1 = p->d1hh * (exp(-2.0 * p->ahh * xbh) - 2.0 * exp(-p->ahh * xbh));
        e3 = p->d3hh * (exp(-2.0 * p->ahh * xbh) + 2.0 * exp(-p->ahh * xbh));
        vqbh[i] = (e1 + e3) * 0.5;
        vjbh[i] = (e1 - e3) * 0.5;
        vq[i] = vqch[i] + vqcb + vqbh[i];
        const double u = vjch[i] - vjcb, v = vjcb - vjbh[i], s = vjbh[i] - vjch[i];
        vj[i] = -sqrt((u * u + v * v + s * s) * 0.5);
        vstr = vstr + vq[i] + vj[i];
    }
    for (int ind = 0; ind < 3; ++ind) {
        achdc[ind] = dumach * (w->tch[0][ind] / w->rch[0] + w->tch[1][ind] / w->rch[1]
                               + w->tch[2][ind] / w->rch[2] + w->tch[3][ind] / w->rch[3]) / 4.0;
        for (int i = 0; i < 4; ++i) achdh[i][ind] = -dumach * w->tch[i][ind] / w->rch[i] / 4.0;
    }
    const double dumqcb = -p->acb * ((p->d1cb + p->d3cb) * exp(-2.0 * p->acb * xcb)
                                     - (p->d1cb - p->d3cb) * exp(-p->acb * xcb)) / w->rcb;
    for (int i = 0; i < 4; ++i) {
        const double xch = w->rch[i] - p->r0ch, xbh = w->rbh[i] - p->r0hh;
        const double dumqbh = -p->ahh * ((p->d1hh + p->d3hh) * exp(-2.0 * p->ahh * xbh)
                                         - (p->d1hh - p->d3hh) * exp(-p->ahh * xbh)) / w->rbh[i];
        const double factj = 0.5 / vj[i];
        const double dumjcb = -p->acb * ((p->d1cb - p->d3cb) * exp(-2.0 * p->acb * xcb)
                                         - (p->d1cb + p->d3cb) * exp(-p->acb * xcb)) * factj / w->rcb;
        const double dumjbh = -p->ahh * ((p->d1hh - p->d3hh) * exp(-2.0 * p->ahh * xbh)
                                         - (p->d1hh + p->d3hh) * exp(-p->ahh * xbh)) * factj / w->rbh[i];
        const double gq = (p->d1ch + p->d3ch) * exp(-2.0 * ach * xch) - (p->d1ch - p->d3ch) * exp(-ach * xch);
        const double gj = (p->d1ch - p->d3ch) * exp(-2.0 * ach * xch) - (p->d1ch + p->d3ch) * exp(-ach * xch);
        const double u = vjch[i] - vjcb, v = vjcb - vjbh[i], s = vjbh[i] - vjch[i];
        for (int ind = 0; ind < 3; ++ind) {
            const double tcb = w->tcb[ind], tbh = w->tbh[i][ind], tch = w->tch[i][ind];
            const int nb = w->nhb[ind], nc = w->nc[ind], nh = w->nh[i][ind];
            pd[nb] = pd[nb] - tcb * dumqcb + tbh * dumqbh + u * (dumjcb * tcb)
                     + v * (-dumjcb * tcb - dumjbh * tbh) + s * dumjbh * tbh;
            const double dumqch = -(ach * tch / w->rch[i] + achdc[ind] * xch) * gq;
            pd[nc] = pd[nc] + dumqch + tcb * dumqcb;
            const double dumqhi = (ach * tch / w->rch[i] - achdh[i][ind] * xch) * gq;
            pd[nh] = pd[nh] + dumqhi - tbh * dumqbh;
            const double dumjch = -(ach * tch / w->rch[i] + achdc[ind] * xch) * gj * factj;
            pd[nc] = pd[nc] + u * (dumjch - dumjcb * tcb) + v * dumjcb * tcb - s * dumjch;
            const double dumjhi = (ach * tch / w->rch[i] - achdh[i][ind] * xch) * gj * factj;
            pd[nh] = pd[nh] + u * dumjhi + v * dumjbh * tbh + s * (-dumjbh * tbh - dumjhi);
            for (int k = 1; k <= 3; ++k) {
                const int j = (i + k) & 3;
                const double dumqhj = -achdh[j][ind] * xch * gq;
                const double dumjhj = -achdh[j][ind] * xch * gj * factj;
                const int nhj = w->nh[j][ind];
                pd[nhj] = pd[nhj] + dumqhj + u * dumjhj - s * dumjhj;
            }
        }
    }
    return vstr;
}

/* opforce_ch4h: out-of-plane force constants switched by s3 (src/egrad_ch4h.f:1240-1345). */
static void opforce(ch4h_ws *w, const ch4h_par *p)
{
    const double *s3 = w->s3, *ds3 = w->ds3;
    for (int i = 0; i < 4; ++i) {
        const int a = (i + 1) & 3, b = (i + 2) & 3, c = (i + 3) & 3;
        const double sw = (1.0 - s3[i]) * s3[a] * s3[b] * s3[c];
        double dsw[4];
        dsw[i] = -ds3[i] * s3[a] * s3[b] * s3[c];
        dsw[a] = (1.0 - s3[i]) * ds3[a] * s3[b] * s3[c];
        dsw[b] = (1.0 - s3[i]) * s3[a] * ds3[b] * s3[c];
        dsw[c] = (1.0 - s3[i]) * s3[a] * s3[b] * ds3[c];
        w->fdelta[i] = sw * p->fch3;
        w->hdelta[i] = sw * p->hch3;
        for (int j = 0; j < 4; ++j) {
            w->dfdelta[i][j] = dsw[j] * p->fch3;
            w->dhdelta[i][j] = dsw[j] * p->hch3;
        }
    }
}

/* calcdelta_ch4h: out-of-plane angles delta(i,m), m in (j,k,l), and their
 * Cartesian derivatives (src/egrad_ch4h.f:986-1236). Indices are 0-based. */
static void calcdelta(ch4h_ws *w, int i, int j, int k, int l, double *sum2, double *sum4)
{
    double *pd = w->pdot;
    double delta[4], a[3], b[3], axb[3], c[4][3], argd[4], daxb[4][3][3], cdot[4][3][3], t2[3];
    const int in[3] = { j, k, l };
    *sum2 = 0.0;
    *sum4 = 0.0;
    for (int ind = 0; ind < 3; ++ind) {
        a[ind] = w->q[w->nh[k][ind]] - w->q[w->nh[j][ind]];
        b[ind] = w->q[w->nh[l][ind]] - w->q[w->nh[j][ind]];
    }
    axb[0] = a[1] * b[2] - a[2] * b[1];
    axb[1] = a[2] * b[0] - a[0] * b[2];
    axb[2] = a[0] * b[1] - a[1] * b[0];
    double norma = axb[0] * axb[0] + axb[1] * axb[1] + axb[2] * axb[2];
    norma = sqrt(norma);
    for (int ii = 0; ii < 3; ++ii)
        for (int ind = 0; ind < 3; ++ind) c[in[ii]][ind] = -w->tch[in[ii]][ind] / w->rch[in[ii]];
    for (int ii = 0; ii < 3; ++ii) {
        const int m = in[ii];
        argd[m] = axb[0] * c[m][0] + axb[1] * c[m][1] + axb[2] * c[m][2];
        argd[m] = argd[m] / norma;
        delta[m] = acos(argd[m]) - w->theta0[i][m];
        *sum2 = *sum2 + delta[m] * delta[m];
        const double d2 = delta[m] * delta[m];
        *sum4 = *sum4 + d2 * d2;
    }
    /* d(a x b)/d h_j, d h_k, d h_l: daxb[atom][dir][component] */
    daxb[j][0][0] = 0.0;          daxb[j][0][1] = b[2] - a[2];  daxb[j][0][2] = -b[1] + a[1];
    daxb[j][1][0] = -b[2] + a[2]; daxb[j][1][1] = 0.0;          daxb[j][1][2] = b[0] - a[0];
    daxb[j][2][0] = b[1] - a[1];  daxb[j][2][1] = -b[0] + a[0]; daxb[j][2][2] = 0.0;
    daxb[k][0][0] = 0.0;   daxb[k][0][1] = -b[2]; daxb[k][0][2] = b[1];
    daxb[k][1][0] = b[2];  daxb[k][1][1] = 0.0;   daxb[k][1][2] = -b[0];
    daxb[k][2][0] = -b[1]; daxb[k][2][1] = b[0];  daxb[k][2][2] = 0.0;
    daxb[l][0][0] = 0.0;   daxb[l][0][1] = a[2];  daxb[l][0][2] = -a[1];
    daxb[l][1][0] = -a[2]; daxb[l][1][1] = 0.0;   daxb[l][1][2] = a[0];
    daxb[l][2][0] = a[1];  daxb[l][2][1] = -a[0]; daxb[l][2][2] = 0.0;
    /* dc(m)_x/dh(m)_y = delta_xy/r + t_y c_x / r^2: cdot[m][x][y] */
    for (int ii = 0; ii < 3; ++ii) {
        const int m = in[ii];
        const double r = w->rch[m], r2 = r * r;
        for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y)
                cdot[m][x][y] = (x == y) ? 1.0 / r + w->tch[m][y] * c[m][x] / r2
                                         : w->tch[m][y] * c[m][x] / r2;
    }
    const double norma3 = norma * norma * norma;
    for (int ii = 0; ii < 3; ++ii) {
        const int m = in[ii];
        const double dm = delta[m], dm3 = dm * dm * dm;
        for (int ind = 0; ind < 3; ++ind) {
            double deldot = -w->dtheta0[i][m][i];
            deldot = -deldot * w->tch[i][ind] / w->rch[i];
            pd[w->nh[i][ind]] = pd[w->nh[i][ind]] + 2.0 * w->fdelta[i] * dm * deldot
                                + 4.0 * w->hdelta[i] * dm3 * deldot;
            deldot = -deldot;
            pd[w->nc[ind]] = pd[w->nc[ind]] + 2.0 * w->fdelta[i] * dm * deldot
                             + 4.0 * w->hdelta[i] * dm3 * deldot;
            for (int jj = 0; jj < 3; ++jj) {
                const int n = in[jj];
                double at1 = axb[0] * daxb[n][ind][0] + axb[1] * daxb[n][ind][1] + axb[2] * daxb[n][ind][2];
                at1 = at1 / norma3;
                t2[0] = daxb[n][ind][0] / norma - at1 * axb[0];
                t2[1] = daxb[n][ind][1] / norma - at1 * axb[1];
                t2[2] = daxb[n][ind][2] / norma - at1 * axb[2];
                const double at3 = t2[0] * c[m][0] + t2[1] * c[m][1] + t2[2] * c[m][2];
                double at4 = 0.0;
                if (ii == jj) {
                    at4 = axb[0] * cdot[m][0][ind] + axb[1] * cdot[m][1][ind] + axb[2] * cdot[m][2][ind];
                    at4 = at4 / norma;
                }
                double at5 = -w->dtheta0[i][m][n];
                at5 = -at5 * w->tch[n][ind] / w->rch[n];
                const double sq = sqrt(1.0 - argd[m] * argd[m]);
                deldot = at3 + at4;
                deldot = -1.0 / sq * deldot;
                deldot = deldot + at5;
                pd[w->nh[n][ind]] = pd[w->nh[n][ind]] + 2.0 * w->fdelta[i] * dm * deldot
                                    + 4.0 * w->hdelta[i] * dm3 * deldot;
                deldot = 1.0 / sq * at4;
                deldot = deldot - at5;
                pd[w->nc[ind]] = pd[w->nc[ind]] + 2.0 * w->fdelta[i] * dm * deldot
                                 + 4.0 * w->hdelta[i] * dm3 * deldot;
            }
        }
    }
}

/* opbend_ch4h (src/egrad_ch4h.f:713-860) */
static double opbend(ch4h_ws *w, const ch4h_par *p)
{
    double sumd2[4], sumd4[4], a[3], b[3], axb[3], c[3];
    double vop = 0.0;
    opforce(w, p);
    for (int i = 0; i < 4; ++i) {
        const int j = (i + 1) & 3;
        int k = (j + 1) & 3, l = (k + 1) & 3;
        const int in[3] = { j, k, l };
        for (int ind = 0; ind < 3; ++ind) {
            a[ind] = w->q[w->nh[k][ind]] - w->q[w->nh[j][ind]];
            b[ind] = w->q[w->nh[l][ind]] - w->q[w->nh[j][ind]];
        }
        axb[0] = a[1] * b[2] - a[2] * b[1];
        axb[1] = a[2] * b[0] - a[0] * b[2];
        axb[2] = a[0] * b[1] - a[1] * b[0];
        double norma = axb[0] * axb[0] + axb[1] * axb[1] + axb[2] * axb[2];
        norma = sqrt(norma);
        /* the reference swaps k and l once for every positive projection */
        for (int ii = 0; ii < 3; ++ii) {
            const int m = in[ii];
            for (int ind = 0; ind < 3; ++ind) c[ind] = -w->tch[m][ind] / w->rch[m];
            double argd = axb[0] * c[0] + axb[1] * c[1] + axb[2] * c[2];
            argd = argd / norma;
            if (argd > 0.0) {
                const int t = k;
                k = l;
                l = t;
            }
        }
        double s2, s4;
        calcdelta(w, i, j, k, l, &s2, &s4);
        sumd2[i] = s2;
        sumd4[i] = s4;
        vop = vop + w->fdelta[i] * sumd2[i] + w->hdelta[i] * sumd4[i];
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            const double ddr = w->dfdelta[i][j] * sumd2[i] + w->dhdelta[i][j] * sumd4[i];
            for (int ind = 0; ind < 3; ++ind) {
                w->pdot[w->nh[j][ind]] = w->pdot[w->nh[j][ind]] - w->tch[j][ind] * ddr / w->rch[j];
                w->pdot[w->nc[ind]] = w->pdot[w->nc[ind]] + w->tch[j][ind] * ddr / w->rch[j];
            }
        }
    return vop;
}

/* ipforce_ch4h: in-plane force constants (src/egrad_ch4h.f:1350-1540). */
static void ipforce(ch4h_ws *w, const ch4h_par *p)
{
    double dfk0[4][4][4], df1dc[4], df1dh[4];
    const double f0 = p->fkinf + p->ak, f2 = p->fkinf;
    const double *s1 = w->s1, *ds1 = w->ds1, *s2 = w->s2, *ds2 = w->ds2;
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 4; ++j) {
            int o[2], n = 0;                      /* the complementary pair, ascending */
            for (int m = 0; m < 4; ++m)
                if (m != i && m != j) o[n++] = m;
            w->fk0[i][j] = f0 + f0 * (s1[i] * s1[j] - 1.0) + (f0 - f2) * (s2[o[0]] * s2[o[1]] - 1.0);
            dfk0[i][j][i] = f0 * ds1[i] * s1[j];
            dfk0[i][j][j] = f0 * s1[i] * ds1[j];
            dfk0[i][j][o[0]] = (f0 - f2) * ds2[o[0]] * s2[o[1]];
            dfk0[i][j][o[1]] = (f0 - f2) * s2[o[0]] * ds2[o[1]];
        }
    for (int i = 0; i < 4; ++i) {
        const double rbh = w->rbh[i], xch = w->rch[i] - p->r0ch;
        const double arga1 = p->aa1 * rbh * rbh;
        const double arga2 = p->aa4 * (rbh - p->r0hh) * (rbh - p->r0hh);
        const double a1 = 1.0 - exp(-arga1);
        const double a2 = p->aa2 + p->aa3 * exp(-arga2);
        w->f1[i] = a1 * exp(-a2 * (xch * xch));
        const double duma1 = 2.0 * p->aa1 * rbh * exp(-arga1);
        const double duma2 = -2.0 * p->aa3 * p->aa4 * (rbh - p->r0hh) * exp(-arga2);
        df1dc[i] = -2.0 * xch * a1 * a2 * exp(-a2 * (xch * xch));
        df1dh[i] = duma1 * exp(-a2 * (xch * xch)) - duma2 * (xch * xch) * a1 * exp(-a2 * (xch * xch));
    }
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 4; ++j)
            for (int k = 0; k < 4; ++k) {
                double v = dfk0[i][j][k] * w->f1[i] * w->f1[j];
                if (k == i) v = v + w->fk0[i][j] * df1dc[i] * w->f1[j];
                if (k == j) v = v + w->fk0[i][j] * w->f1[i] * df1dc[j];
                w->dfdc[i][j][k] = v;
                if (k == i) w->dfdh[i][j][k] = w->fk0[i][j] * df1dh[i] * w->f1[j];
                else if (k == j) w->dfdh[i][j][k] = w->fk0[i][j] * w->f1[i] * df1dh[j];
                else w->dfdh[i][j][k] = 0.0;
            }
}

/* ipbend_ch4h (src/egrad_ch4h.f:865-982) */
static double ipbend(ch4h_ws *w, const ch4h_par *p)
{
    double *pd = w->pdot;
    double vip = 0.0;
    ipforce(w, p);
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 4; ++j) {
            double costh = w->tch[i][0] * w->tch[j][0] + w->tch[i][1] * w->tch[j][1] + w->tch[i][2] * w->tch[j][2];
            costh = costh / w->rch[i] / w->rch[j];
            const double theta = acos(costh);
            const double dth = theta - w->theta0[i][j];
            const double fff = w->fk0[i][j] * w->f1[i] * w->f1[j];
            vip = vip + 0.5 * w->fk0[i][j] * w->f1[i] * w->f1[j] * (dth * dth);
            const double termth = -1.0 / sqrt(1.0 - costh * costh);
            for (int ind = 0; ind < 3; ++ind) {
                double dthi = -w->tch[j][ind] / w->rch[i] / w->rch[j] + costh * w->tch[i][ind] / w->rch[i] / w->rch[i];
                dthi = dthi * termth;
                double dthj = -w->tch[i][ind] / w->rch[i] / w->rch[j] + costh * w->tch[j][ind] / w->rch[j] / w->rch[j];
                dthj = dthj * termth;
                const double dthc = -(dthi + dthj);
                pd[w->nh[i][ind]] = pd[w->nh[i][ind]] + fff * dthi * dth;
                pd[w->nh[j][ind]] = pd[w->nh[j][ind]] + fff * dthj * dth;
                pd[w->nc[ind]] = pd[w->nc[ind]] + fff * dthc * dth;
                for (int k = 0; k < 4; ++k) {
                    const double dth0k = -w->dtheta0[i][j][k] * w->tch[k][ind] / w->rch[k];
                    const double dth0c = -dth0k;
                    pd[w->nh[k][ind]] = pd[w->nh[k][ind]]
                                        - 0.5 * w->tch[k][ind] * w->dfdc[i][j][k] * (dth * dth) / w->rch[k]
                                        - 0.5 * w->tbh[k][ind] * w->dfdh[i][j][k] * (dth * dth) / w->rbh[k]
                                        - fff * dth0k * dth;
                    pd[w->nc[ind]] = pd[w->nc[ind]]
                                     + 0.5 * w->tch[k][ind] * w->dfdc[i][j][k] * (dth * dth) / w->rch[k]
                                     - fff * dth0c * dth;
                    pd[w->nhb[ind]] = pd[w->nhb[ind]]
                                      + 0.5 * w->tbh[k][ind] * w->dfdh[i][j][k] * (dth * dth) / w->rbh[k];
                }
            }
        }
    return vip;
}

/* pot_ch4h for one bead: x(3,6) bohr -> energy (Eh), g(3,6) (Eh/bohr). */
void oracle_ch4h_pot(const double *x, double *energy, double *g)
{
    const ch4h_par p = make_par();
    ch4h_ws w;
    set_index(&w);
    for (int i = 0; i < 18; ++i) {
        w.q[i] = x[i] * 0.52918;
        w.pdot[i] = 0.0;
    }
    coorden(&w);
    switchf(&w, &p);
    refangles(&w);
    const double vstr = stretch(&w, &p);
    const double vop = opbend(&w, &p);
    const double vip = ipbend(&w, &p);
    double en = vstr + vop + vip;
    *energy = en * 0.03812;
    for (int i = 0; i < 18; ++i) g[i] = w.pdot[i] * 0.0201723;
}

/* Bead-batched plugin entry, same contract as egrad_ch4h (src/egrad_ch4h.f:74-129). */
void oracle_egrad_ch4h(const double *q, int natoms, int nbeads, double *v, double *dvdq)
{
    for (int k = 0; k < nbeads; ++k)
        oracle_ch4h_pot(q + (size_t)k * 3 * natoms, &v[k], dvdq + (size_t)k * 3 * natoms);
}
