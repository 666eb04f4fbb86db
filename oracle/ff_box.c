/*
 * oracle/ff_box.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * QMDFF-form force field of a periodic box of small molecules, in plain C
 * double precision, all pairs:
 *   bonds    De (1 - exp(-a (r - r0)))^2                         (Morse)
 *   angles   k (theta - theta0)^2           (the k(theta-theta0)^2 form of
 *                                            ff_eg, src/ff_eg.f90:167-299)
 *   Coulomb  q_i q_j (erfc(a r)/r - zp (r - rc)), r < rc          (Zahn damped
 *            shifted force, src/ff_nonb.f90:376-394, src/set_periodic.f90:93-97,
 *            with the energy-consistent gradient)
 *   dispersion -(C6/(r^6 + R0^6) + sr42 C6/(r^8 + R0^8)), r < rc    (BJ-damped
 *            D3 form of ff_nonb, src/ff_nonb.f90:198-290)
 *   repulsion zab exp(-alpha r)/r, r < rc                          (ff_nonb's
 *            Pauli repulsion, src/ff_nonb.f90:292-300)
 * Nonbonded terms act between atoms of different molecules only (the
 * reference's multi-molecule loop, src/ff_nonb.f90:198-206); distances use the
 * minimum image (box_image, src/box_image.f90:36-64).
 *
 * Parameters theta[10] = De, a, r0, k, theta0, q_O, q_H, C6_OO, C6_OH, C6_HH.
 * Returns E; optionally dE/dx, dE/dtheta, and for a direction v (natoms x 3)
 * d(v . dE/dx)/dtheta, the contraction force matching needs.
 * Pinned by central differences of E in x and in theta and of (v . dE/dx)
 * in theta, and by closed forms of single terms (tests/test_oracle_ff.py).
 */
#include <math.h>
#include <stddef.h>
#include "oracle.h"

static double min_image(double d, double box) { return d - box * rint(d / box); }

double oracle_ff_eval(const oracle_ff *f, const double *th, const double *x, double *g, double *dth,
                      const double *v, double *dvth)
{
    const int n = f->natoms;
    double e = 0.0;
    if (g)
        for (int i = 0; i < 3 * n; ++i) g[i] = 0.0;
    if (dth)
        for (int i = 0; i < ORACLE_FF_NPAR; ++i) dth[i] = 0.0;
    if (dvth)
        for (int i = 0; i < ORACLE_FF_NPAR; ++i) dvth[i] = 0.0;
    const double De = th[0], aa = th[1], r0 = th[2], kk = th[3], t0 = th[4];
    /* ---- bonds ---- */
    for (int b = 0; b < f->nbonds; ++b) {
        const int i = f->bonds[2 * b], j = f->bonds[2 * b + 1];
        double d[3], r = 0.0;
        for (int c = 0; c < 3; ++c) {
            d[c] = min_image(x[3 * i + c] - x[3 * j + c], f->box);
            r += d[c] * d[c];
        }
        r = sqrt(r);
        const double xx = r - r0, ex = exp(-aa * xx), om = 1.0 - ex;
        e += De * om * om;
        const double fp = 2.0 * De * aa * ex * om;            /* dE/dr */
        double proj = 0.0;
        for (int c = 0; c < 3; ++c) {
            if (g) {
                g[3 * i + c] += fp * d[c] / r;
                g[3 * j + c] -= fp * d[c] / r;
            }
            if (v) proj += (v[3 * i + c] - v[3 * j + c]) * d[c] / r;
        }
        if (dth) {
            dth[0] += om * om;
            dth[1] += 2.0 * De * om * xx * ex;
            dth[2] += -2.0 * De * om * aa * ex;
        }
        if (dvth) {
            dvth[0] += 2.0 * aa * ex * om * proj;
            dvth[1] += 2.0 * De * ex * (om * (1.0 - aa * xx) + aa * xx * ex) * proj;
            dvth[2] += 2.0 * De * aa * aa * ex * (1.0 - 2.0 * ex) * proj;
        }
    }
    /* ---- angles (vertex in the middle) ---- */
    for (int a = 0; a < f->nangles; ++a) {
        const int i = f->angles[3 * a], j = f->angles[3 * a + 1], k = f->angles[3 * a + 2];
        double u[3], w[3], ru = 0.0, rw = 0.0, dot = 0.0;
        for (int c = 0; c < 3; ++c) {
            u[c] = min_image(x[3 * i + c] - x[3 * j + c], f->box);
            w[c] = min_image(x[3 * k + c] - x[3 * j + c], f->box);
            ru += u[c] * u[c];
            rw += w[c] * w[c];
            dot += u[c] * w[c];
        }
        ru = sqrt(ru);
        rw = sqrt(rw);
        const double cs = dot / (ru * rw);
        const double th_ = acos(cs);
        const double dt = th_ - t0;
        e += kk * dt * dt;
        const double de = 2.0 * kk * dt;                       /* dE/dtheta */
        const double s = -1.0 / sqrt(1.0 - cs * cs);           /* dtheta/dcos */
        double gi[3], gk[3], proj = 0.0;
        for (int c = 0; c < 3; ++c) {
            gi[c] = s * (w[c] / rw - cs * u[c] / ru) / ru;     /* dtheta/dx_i */
            gk[c] = s * (u[c] / ru - cs * w[c] / rw) / rw;     /* dtheta/dx_k */
            if (v) proj += v[3 * i + c] * gi[c] + v[3 * k + c] * gk[c] - v[3 * j + c] * (gi[c] + gk[c]);
        }
        if (g)
            for (int c = 0; c < 3; ++c) {
                g[3 * i + c] += de * gi[c];
                g[3 * k + c] += de * gk[c];
                g[3 * j + c] -= de * (gi[c] + gk[c]);
            }
        if (dth) {
            dth[3] += dt * dt;
            dth[4] += -2.0 * kk * dt;
        }
        if (dvth) {
            dvth[3] += 2.0 * dt * proj;
            dvth[4] += -2.0 * kk * proj;
        }
    }
    /* ---- nonbonded, intermolecular pairs within the cutoff ---- */
    const double rc = f->rcut, za = f->zahn_a, zp = f->zahn_par;
    const double spi = 1.7724538509055160;
    for (int i = 0; i < n - 1; ++i)
        for (int j = i + 1; j < n; ++j) {
            if (f->mol[i] == f->mol[j]) continue;
            double d[3], r2 = 0.0;
            for (int c = 0; c < 3; ++c) {
                d[c] = min_image(x[3 * i + c] - x[3 * j + c], f->box);
                r2 += d[c] * d[c];
            }
            const double r = sqrt(r2);
            if (r > rc) continue;
            const int ti = f->type[i], tj = f->type[j];
            const double qi = th[5 + ti], qj = th[5 + tj];
            const int cidx = 7 + ti + tj;                      /* OO -> 7, OH -> 8, HH -> 9 */
            const double c6 = th[cidx];
            /* Coulomb */
            const double erfcv = erfc(za * r);
            const double phi = erfcv / r - zp * (r - rc);
            const double dphi = -erfcv / r2 - 2.0 * za / spi * exp(-za * za * r2) / r - zp;
            /* dispersion */
            const double R0 = f->r0bj[ti][tj], sr = f->sr42[ti][tj];
            const double r4 = r2 * r2, r6 = r4 * r2, R06 = R0 * R0 * R0 * R0 * R0 * R0;
            const double t6 = r6 + R06, t8 = r6 * r2 + R06 * R0 * R0;
            const double dispu = -(1.0 / t6 + sr / t8);                       /* E_disp / C6 */
            const double ddispu = 6.0 * r4 * r / (t6 * t6) + sr * 8.0 * r6 * r / (t8 * t8);   /* d/dr */
            /* repulsion */
            const double za_ = f->zrep[ti][tj], al = f->arep[ti][tj];
            const double rep = za_ * exp(-al * r) / r;
            const double drep = -rep * (al + 1.0 / r);
            e += qi * qj * phi + c6 * dispu + rep;
            const double fp = qi * qj * dphi + c6 * ddispu + drep;
            double proj = 0.0;
            for (int c = 0; c < 3; ++c) {
                if (g) {
                    g[3 * i + c] += fp * d[c] / r;
                    g[3 * j + c] -= fp * d[c] / r;
                }
                if (v) proj += (v[3 * i + c] - v[3 * j + c]) * d[c] / r;
            }
            if (dth) {
                dth[5 + ti] += qj * phi;
                dth[5 + tj] += qi * phi;
                dth[cidx] += dispu;
            }
            if (dvth) {
                dvth[5 + ti] += qj * dphi * proj;
                dvth[5 + tj] += qi * dphi * proj;
                dvth[cidx] += ddispu * proj;
            }
        }
    return e;
}
