/*
 * oracle/rpmd.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * CPU restatement of one RPMD velocity-Verlet step of the reference
 * (src/verlet.f90:65-1308) and of the routines it calls on this path:
 * rfft/irfft (src/rfft.f90, src/irfft.f90), get_centroid, transrot
 * (src/transrot.f90 + src/invert.f90), calc_xi BIMOLEC (src/calc_xi.f90),
 * calc_com, constrain_q, constrain_p, andersen and the recrossing child
 * flux (src/recross_serial.f90:189-233).  Arrays use the reference layout
 * x(3,natoms,nbeads) column-major: x[(k*natoms + i)*3 + d].
 *
 * Parity pins (tests/test_oracle_rpmd.py): the cosine transform against
 * numpy's FFT; the free ring-polymer step against its closed form (exact
 * harmonic propagator, C.C = (I+R)/2); energy conservation of the full
 * ring-polymer Hamiltonian; calc_xi gradients against central
 * differences; the constraint against its defining equations; Philox
 * against the published known-answer vectors.
 */
#include <math.h>
#include <string.h>
#include "oracle.h"

#define IDX(k, i, d, na) (((size_t)(k) * (na) + (i)) * 3 + (d))

/* rfft/irfft: x <- sqrt(1/N) * Re(sum_j x_j exp(-2 pi i k j / N)).  FFTW is
 * replaced by the direct sum it computes (FFTW_FORWARD, real part). */
void oracle_rfft(double *x, int n)
{
    double y[ORACLE_MAX_BEADS];
    const double pi = 3.141592653589793;
    const double factor = sqrt(1.0 / n);
    for (int k = 0; k < n; ++k) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) {
            const long m = ((long)k * j) % n;
            s += x[j] * cos(2.0 * pi * (double)m / n);
        }
        y[k] = factor * s;
    }
    memcpy(x, y, sizeof(double) * n);
}

void oracle_forces(const oracle_sys *s, const double *q, double *v, double *g)
{
    if (s->surface == ORACLE_SURF_H3)
        oracle_egrad_h3(q, s->natoms, s->nbeads, v, g);
    else if (s->surface == ORACLE_SURF_OH3)
        oracle_egrad_oh3(q, s->natoms, s->nbeads, v, g);
    else if (s->surface == ORACLE_SURF_BRH2)
        oracle_egrad_brh2(q, s->natoms, s->nbeads, v, g);
    else if (s->surface == ORACLE_SURF_O3)
        oracle_egrad_o3(q, s->natoms, s->nbeads, v, g);
    else
        oracle_egrad_ch4h(q, s->natoms, s->nbeads, v, g);
}

/* Free ring-polymer step in normal-mode space (src/verlet.f90:377-463). */
void oracle_free_rp(const oracle_sys *s, double *q, double *p, double beta)
{
    const int na = s->natoms, nb = s->nbeads;
    const double pi = 3.141592653589793;
    double poly[4][ORACLE_MAX_BEADS];
    double vp[ORACLE_MAX_BEADS], vq[ORACLE_MAX_BEADS];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < na; ++j) {
            for (int k = 0; k < nb; ++k) {
                vp[k] = p[IDX(k, j, i, na)];
                vq[k] = q[IDX(k, j, i, na)];
            }
            oracle_rfft(vp, nb);
            oracle_rfft(vq, nb);
            for (int k = 0; k < nb; ++k) {
                p[IDX(k, j, i, na)] = vp[k];
                q[IDX(k, j, i, na)] = vq[k];
            }
        }
    for (int j = 0; j < na; ++j) {
        poly[0][0] = 1.0;
        poly[1][0] = 0.0;
        poly[2][0] = s->dt / s->mass[j];
        poly[3][0] = 1.0;
        if (nb > 1) {
            const double beta_n = beta / nb;
            const double twown = 2.0 / beta_n;
            const double pi_n = pi / nb;
            for (int k = 1; k <= nb / 2; ++k) {
                const double wk = twown * sin(k * pi_n);
                const double wt = wk * s->dt;
                const double wm = wk * s->mass[j];
                const double cw = cos(wt), sw = sin(wt);
                poly[0][k] = cw;
                poly[1][k] = -wm * sw;
                poly[2][k] = sw / wm;
                poly[3][k] = cw;
            }
            for (int k = 1; k <= (nb - 1) / 2; ++k)
                for (int c = 0; c < 4; ++c) poly[c][nb - k] = poly[c][k];
        }
        for (int k = 0; k < nb; ++k)
            for (int i = 0; i < 3; ++i) {
                const size_t o = IDX(k, j, i, na);
                const double pn = p[o] * poly[0][k] + q[o] * poly[1][k];
                q[o] = p[o] * poly[2][k] + q[o] * poly[3][k];
                p[o] = pn;
            }
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < na; ++j) {
            for (int k = 0; k < nb; ++k) {
                vp[k] = p[IDX(k, j, i, na)];
                vq[k] = q[IDX(k, j, i, na)];
            }
            oracle_rfft(vp, nb);
            oracle_rfft(vq, nb);
            for (int k = 0; k < nb; ++k) {
                p[IDX(k, j, i, na)] = vp[k];
                q[IDX(k, j, i, na)] = vq[k];
            }
        }
}

void oracle_centroid(const oracle_sys *s, const double *q, double *c)
{
    const int na = s->natoms, nb = s->nbeads;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < na; ++j) {
            double acc = 0.0;
            for (int k = 0; k < nb; ++k) acc = acc + q[IDX(k, j, i, na)];
            c[j * 3 + i] = acc / nb;
        }
}

/* Gauss-Jordan inversion with full pivoting, column-major a(n,n) (src/invert.f90). */
static void invert3(double a[3][3])   /* a[col][row] == a(row,col) */
{
    int ipivot[3] = { 0, 0, 0 }, indxr[3], indxc[3];
#define A(r, c) a[c][r]
    for (int i = 0; i < 3; ++i) {
        double big = 0.0;
        int irow = 0, icol = 0;
        for (int j = 0; j < 3; ++j)
            if (ipivot[j] != 1)
                for (int k = 0; k < 3; ++k)
                    if (ipivot[k] == 0 && fabs(A(j, k)) >= big) {
                        big = fabs(A(j, k));
                        irow = j;
                        icol = k;
                    }
        ipivot[icol] += 1;
        if (irow != icol)
            for (int j = 0; j < 3; ++j) {
                const double t = A(irow, j);
                A(irow, j) = A(icol, j);
                A(icol, j) = t;
            }
        indxr[i] = irow;
        indxc[i] = icol;
        const double pivot = A(icol, icol);
        A(icol, icol) = 1.0;
        for (int j = 0; j < 3; ++j) A(icol, j) = A(icol, j) / pivot;
        for (int j = 0; j < 3; ++j)
            if (j != icol) {
                const double t = A(j, icol);
                A(j, icol) = 0.0;
                for (int k = 0; k < 3; ++k) A(j, k) = A(j, k) - A(icol, k) * t;
            }
    }
    for (int i = 2; i >= 0; --i)
        if (indxr[i] != indxc[i])
            for (int k = 0; k < 3; ++k) {
                const double t = A(k, indxr[i]);
                A(k, indxr[i]) = A(k, indxc[i]);
                A(k, indxc[i]) = t;
            }
#undef A
}

/* transrot (src/transrot.f90:36-237), including its total mass n^2 * M. */
void oracle_transrot(const oracle_sys *s, const double *q, double *p)
{
    const int na = s->natoms, nb = s->nbeads;
    double vel[3 * ORACLE_MAX_ATOMS * ORACLE_MAX_BEADS];
    for (int k = 0; k < nb; ++k)
        for (int i = 0; i < na; ++i)
            for (int d = 0; d < 3; ++d) vel[IDX(k, i, d, na)] = p[IDX(k, i, d, na)] / s->mass[i];
    double totmass = 0.0, vtot[3] = { 0, 0, 0 };
    for (int i = 0; i < na; ++i) {
        const double w = s->mass[i];
        for (int k = 0; k < nb; ++k) {
            totmass = totmass + w;
            for (int j = 0; j < 3; ++j) vtot[j] = vtot[j] + vel[IDX(k, i, j, na)] * w;
        }
    }
    totmass = totmass * nb;
    for (int j = 0; j < 3; ++j) vtot[j] = vtot[j] / totmass;
    double xt = 0.0, yt = 0.0, zt = 0.0;
    for (int i = 0; i < na; ++i) {
        const double w = s->mass[i];
        for (int k = 0; k < nb; ++k) {
            xt = xt + q[IDX(k, i, 0, na)] * w;
            yt = yt + q[IDX(k, i, 1, na)] * w;
            zt = zt + q[IDX(k, i, 2, na)] * w;
        }
    }
    xt = xt / totmass;
    yt = yt / totmass;
    zt = zt / totmass;
    double mang[3] = { 0, 0, 0 };
    for (int i = 0; i < na; ++i) {
        const double w = s->mass[i];
        for (int k = 0; k < nb; ++k) {
            const double *x = &q[IDX(k, i, 0, na)], *v = &vel[IDX(k, i, 0, na)];
            mang[0] = mang[0] + (x[1] * v[2] - x[2] * v[1]) * w;
            mang[1] = mang[1] + (x[2] * v[0] - x[0] * v[2]) * w;
            mang[2] = mang[2] + (x[0] * v[1] - x[1] * v[0]) * w;
        }
    }
    mang[0] = mang[0] - (yt * vtot[2] - zt * vtot[1]) * totmass;
    mang[1] = mang[1] - (zt * vtot[0] - xt * vtot[2]) * totmass;
    mang[2] = mang[2] - (xt * vtot[1] - yt * vtot[0]) * totmass;
    double xx = 0, xy = 0, xz = 0, yy = 0, yz = 0, zz = 0;
    for (int i = 0; i < na; ++i) {
        const double w = s->mass[i];
        for (int k = 0; k < nb; ++k) {
            const double xd = q[IDX(k, i, 0, na)] - xt, yd = q[IDX(k, i, 1, na)] - yt, zd = q[IDX(k, i, 2, na)] - zt;
            xx = xx + xd * xd * w;
            xy = xy + xd * yd * w;
            xz = xz + xd * zd * w;
            yy = yy + yd * yd * w;
            yz = yz + yd * zd * w;
            zz = zz + zd * zd * w;
        }
    }
    double t[3][3];   /* t[col][row] */
    t[0][0] = yy + zz; t[0][1] = -xy;     t[0][2] = -xz;
    t[1][0] = -xy;     t[1][1] = xx + zz; t[1][2] = -yz;
    t[2][0] = -xz;     t[2][1] = -yz;     t[2][2] = xx + yy;
    if (na <= 2) {
        t[0][0] += 0.000001;
        t[1][1] += 0.000001;
        t[2][2] += 0.000001;
    }
    invert3(t);
    double vang[3];
    for (int i = 0; i < 3; ++i) {
        vang[i] = 0.0;
        for (int j = 0; j < 3; ++j) vang[i] = vang[i] + t[j][i] * mang[j];
    }
    for (int i = 0; i < na; ++i)
        for (int k = 0; k < nb; ++k)
            for (int j = 0; j < 3; ++j) vel[IDX(k, i, j, na)] = vel[IDX(k, i, j, na)] - vtot[j];
    for (int i = 0; i < na; ++i)
        for (int k = 0; k < nb; ++k) {
            double *v = &vel[IDX(k, i, 0, na)];
            const double xd = q[IDX(k, i, 0, na)] - xt, yd = q[IDX(k, i, 1, na)] - yt, zd = q[IDX(k, i, 2, na)] - zt;
            v[0] = v[0] - vang[1] * zd + vang[2] * yd;
            v[1] = v[1] - vang[2] * xd + vang[0] * zd;
            v[2] = v[2] - vang[0] * yd + vang[1] * xd;
        }
    for (int k = 0; k < nb; ++k)
        for (int i = 0; i < na; ++i)
            for (int d = 0; d < 3; ++d) p[IDX(k, i, d, na)] = vel[IDX(k, i, d, na)] * s->mass[i];
}

/* calc_xi BIMOLEC (src/calc_xi.f90:63-282) on coords c(3,natoms); mode 1 or 2.
 * dxi (3,natoms) may be NULL.  Atom indices in rx are 1-based. */
double oracle_calc_xi(const oracle_sys *s, const oracle_rx *rx, const double *c, double *dxi, int mode)
{
    const int na = s->natoms;
    double rf[4][3], rb[4][3], fa[4], ba[4];
    for (int i = 0; i < rx->n_form; ++i) {
        const int a = rx->bond_form[i][0] - 1, b = rx->bond_form[i][1] - 1;
        for (int d = 0; d < 3; ++d) rf[i][d] = c[3 * a + d] - c[3 * b + d];
        fa[i] = sqrt(rf[i][0] * rf[i][0] + rf[i][1] * rf[i][1] + rf[i][2] * rf[i][2]);
    }
    for (int i = 0; i < rx->n_break; ++i) {
        const int a = rx->bond_break[i][0] - 1, b = rx->bond_break[i][1] - 1;
        for (int d = 0; d < 3; ++d) rb[i][d] = c[3 * a + d] - c[3 * b + d];
        ba[i] = sqrt(rb[i][0] * rb[i][0] + rb[i][1] * rb[i][1] + rb[i][2] * rb[i][2]);
    }
    double s1 = 0.0;
    for (int i = 0; i < rx->n_break; ++i) s1 = s1 + (ba[i] - rx->break_ref[i]) / (double)(float)rx->n_break;
    for (int i = 0; i < rx->n_form; ++i) s1 = s1 - (fa[i] - rx->form_ref[i]) / (double)(float)rx->n_form;
    double com[4][3], mreac[4];
    for (int k = 0; k < rx->n_reac; ++k) {
        mreac[k] = 0.0;
        for (int i = 0; i < rx->reac_natoms[k]; ++i) mreac[k] += s->mass[rx->reac_atoms[k][i] - 1];
    }
    for (int k = 0; k < rx->n_reac; ++k) {
        com[k][0] = com[k][1] = com[k][2] = 0.0;
        for (int i = 0; i < rx->reac_natoms[k]; ++i) {
            const int at = rx->reac_atoms[k][i] - 1;
            for (int j = 0; j < 3; ++j) com[k][j] = com[k][j] + s->mass[at] * c[3 * at + j] / mreac[k];
        }
    }
    double red[4][4][3], reds[4][4], s0 = 0.0;
    for (int i = 0; i < rx->n_reac; ++i)
        for (int j = i + 1; j < rx->n_reac; ++j) {
            for (int d = 0; d < 3; ++d) red[i][j][d] = com[j][d] - com[i][d];
            reds[i][j] = sqrt(red[i][j][0] * red[i][j][0] + red[i][j][1] * red[i][j][1] + red[i][j][2] * red[i][j][2]);
            s0 = s0 + (rx->r_inf - reds[i][j]);
        }
    const int terms = (rx->n_reac * rx->n_reac - rx->n_reac) / 2;
    s0 = s0 / (double)(float)terms;
    double xi;
    if (mode == 1) xi = s0 / (s0 - s1);
    else xi = rx->xi_ideal * s1 + (1 - rx->xi_ideal) * s0;
    if (!dxi) return xi;
    double ds1[3 * ORACLE_MAX_ATOMS], ds0[3 * ORACLE_MAX_ATOMS];
    for (int i = 0; i < 3 * na; ++i) ds1[i] = ds0[i] = 0.0;
    for (int i = 0; i < rx->n_form; ++i) {
        const int a = rx->bond_form[i][0] - 1, b = rx->bond_form[i][1] - 1;
        const double rinv = 1.0 / fa[i];
        for (int d = 0; d < 3; ++d) {
            ds1[3 * a + d] = ds1[3 * a + d] - rf[i][d] * rinv / (double)(float)rx->n_form;
            ds1[3 * b + d] = ds1[3 * b + d] + rf[i][d] * rinv / (double)(float)rx->n_form;
        }
    }
    for (int i = 0; i < rx->n_break; ++i) {
        const int a = rx->bond_break[i][0] - 1, b = rx->bond_break[i][1] - 1;
        const double rinv = 1.0 / ba[i];
        for (int d = 0; d < 3; ++d) {
            ds1[3 * a + d] = ds1[3 * a + d] + rb[i][d] * rinv / (double)(float)rx->n_break;
            ds1[3 * b + d] = ds1[3 * b + d] - rb[i][d] * rinv / (double)(float)rx->n_break;
        }
    }
    for (int i = 0; i < rx->n_reac; ++i)
        for (int j = i + 1; j < rx->n_reac; ++j) {
            const double rinv = 1.0 / reds[i][j];
            for (int k = 0; k < rx->reac_natoms[i]; ++k) {
                const int at = rx->reac_atoms[i][k] - 1;
                for (int d = 0; d < 3; ++d)
                    ds0[3 * at + d] = ds0[3 * at + d] + red[i][j][d] * rinv * s->mass[at] / mreac[i] / (double)(float)terms;
            }
            for (int k = 0; k < rx->reac_natoms[j]; ++k) {
                const int at = rx->reac_atoms[j][k] - 1;
                for (int d = 0; d < 3; ++d)
                    ds0[3 * at + d] = ds0[3 * at + d] - red[i][j][d] * rinv * s->mass[at] / mreac[j] / (double)(float)terms;
            }
        }
    for (int i = 0; i < 3 * na; ++i) {
        if (mode == 1) dxi[i] = (s0 * ds1[i] - s1 * ds0[i]) / ((s0 - s1) * (s0 - s1));
        else dxi[i] = rx->xi_ideal * ds1[i] + (1 - rx->xi_ideal) * ds0[i];
    }
    return xi;
}

/* constrain_q (src/constrain_q.f90:35-113): returns 0 on success, 1 if 200
 * Newton iterations were not enough (state unchanged then). */
int oracle_constrain_q(const oracle_sys *s, const oracle_rx *rx, const double *centroid, const double *dxi_act,
                       double *q, double *p)
{
    const int na = s->natoms, nb = s->nbeads;
    double qt[3 * ORACLE_MAX_ATOMS], dxn[3 * ORACLE_MAX_ATOMS];
    double mult = 0.0, coeff = 0.0;
    for (int iter = 1; iter <= 200; ++iter) {
        coeff = mult * s->dt * s->dt / nb;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < na; ++j) qt[3 * j + i] = centroid[3 * j + i] + coeff * dxi_act[3 * j + i] / s->mass[j];
        const double sigma = oracle_calc_xi(s, rx, qt, dxn, 2);
        double dsigma = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < na; ++j)
                dsigma = dsigma + dxn[3 * j + i] * s->dt * s->dt * dxi_act[3 * j + i] / (s->mass[j] * nb);
        const double dx = sigma / dsigma;
        mult = mult - dx;
        if (fabs(dx) < (double)1.0E-8f || fabs(sigma) < (double)1.0E-10f) break;
        if (iter == 200) return 1;
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < na; ++j)
            for (int k = 0; k < nb; ++k) {
                q[IDX(k, j, i, na)] = q[IDX(k, j, i, na)] + coeff / s->mass[j] * dxi_act[3 * j + i];
                p[IDX(k, j, i, na)] = p[IDX(k, j, i, na)] + mult * s->dt / nb * dxi_act[3 * j + i];
            }
    return 0;
}

/* constrain_p (src/constrain_p.f90:36-76). */
void oracle_constrain_p(const oracle_sys *s, const double *dxi_act, double *p)
{
    const int na = s->natoms, nb = s->nbeads;
    double c1 = 0.0, c2 = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < na; ++j)
            for (int k = 0; k < nb; ++k) c1 = c1 + dxi_act[3 * j + i] * p[IDX(k, j, i, na)] / s->mass[j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < na; ++j) c2 = c2 + dxi_act[3 * j + i] * dxi_act[3 * j + i] / s->mass[j];
    const double lam = -c1 / c2 / nb;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < na; ++j)
            for (int k = 0; k < nb; ++k) p[IDX(k, j, i, na)] = p[IDX(k, j, i, na)] + lam * dxi_act[3 * j + i];
}

/* Recrossing flux weight vs/fs of a child (src/recross_serial.f90:189-233). */
double oracle_child_weight(const oracle_sys *s, const oracle_rx *rx, const double *q, const double *p, double beta)
{
    const int na = s->natoms, nb = s->nbeads;
    double c[3 * ORACLE_MAX_ATOMS], dxi[3 * ORACLE_MAX_ATOMS];
    oracle_centroid(s, q, c);
    (void)oracle_calc_xi(s, rx, c, dxi, 2);
    double vs = 0.0;
    for (int l = 0; l < 3; ++l)
        for (int m = 0; m < na; ++m)
            for (int o = 0; o < nb; ++o) vs = vs + dxi[3 * m + l] * p[IDX(o, m, l, na)] / s->mass[m];
    vs = vs / nb;
    double fs = 0.0;
    for (int l = 0; l < 3; ++l)
        for (int m = 0; m < na; ++m) fs = fs + dxi[3 * m + l] * dxi[3 * m + l] / s->mass[m];
    fs = sqrt(fs / (2.0 * 3.141592653589793 * beta));
    return vs / fs;
}

/* andersen (src/andersen.f90:37-74) with counter-based deviates. */
void oracle_andersen(const oracle_sys *s, double *p, double beta, uint64_t seed, uint32_t traj, uint64_t draw)
{
    const int na = s->natoms, nb = s->nbeads;
    const double beta_n = beta / nb;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < na; ++j) {
            const double dp = sqrt(s->mass[j] / beta_n);
            for (int k = 0; k < nb; ++k)
                p[IDX(k, j, i, na)] = oracle_andersen_normal(seed, traj, draw, 3 * j + i, k) * dp;
        }
}

/* One verlet step (src/verlet.f90) for one trajectory.
 *   ensemble 0: dynamic.x NVE (constrain = -1), transrot afterwards if remove_com
 *   ensemble 1: dynamic.x with Andersen every andersen_step steps (thermostat 1)
 *   ensemble 2: recrossing child (constrain = 2): xi of the post-drift centroid
 * g holds dV/dq of the current positions on entry and of the new ones on exit.
 * Returns xi (ensemble 2) or 0. */
double oracle_verlet(const oracle_sys *s, const oracle_rx *rx, double *q, double *p, double *g, double *vbead,
                     double beta, int ensemble, int remove_com, long istep, int andersen_step, uint64_t seed,
                     uint32_t traj)
{
    const int na = s->natoms, nb = s->nbeads;
    const size_t n = (size_t)3 * na * nb;
    for (size_t i = 0; i < n; ++i) p[i] = p[i] - 0.5 * s->dt * g[i];
    if (nb > 1) {
        oracle_free_rp(s, q, p, beta);
    } else {
        /* nbeads == 1: the reference's mode-0 propagator (p, q) -> (p, q + p dt/m) */
        for (int j = 0; j < na; ++j)
            for (int i = 0; i < 3; ++i) {
                const size_t o = IDX(0, j, i, na);
                q[o] = p[o] * (s->dt / s->mass[j]) + q[o] * 1.0;
            }
    }
    double xi = 0.0;
    if (ensemble == 2) {
        double c[3 * ORACLE_MAX_ATOMS];
        oracle_centroid(s, q, c);
        xi = oracle_calc_xi(s, rx, c, NULL, 2);
    }
    oracle_forces(s, q, vbead, g);
    for (size_t i = 0; i < n; ++i) p[i] = p[i] - 0.5 * s->dt * g[i];
    if (ensemble == 1 && andersen_step > 0 && istep % andersen_step == 0)
        oracle_andersen(s, p, beta, seed, traj, (uint64_t)istep);
    if (remove_com && ensemble != 2) oracle_transrot(s, q, p);
    return xi;
}

/* Start of a recrossing parent (mdinit with bias_mode 2, src/mdinit.f90:79-95):
 * the constraint direction dxi is the gradient of the mode-1 reaction
 * coordinate s0/(s0 - s1) at the centroid (umbrella mode 0 -> calc_xi mode 1,
 * src/umbrella.f90:104-106).  The bias k (xi - xi_ideal) dxi that umbrella
 * mode 0 adds to the forces vanishes at a centroid on the dividing surface
 * (mode-1 xi = 1 = xi_ideal there) and is left out. */
/* Hessian of xi (src/calc_xi.f90:283-502) at coords c(3,natoms), mode 1 or
 * 2: h[(3*a+d)*3*natoms + 3*b+e] = d2xi(d,a,e,b), atom-major as c.  s1, s0
 * and their gradients come from oracle_calc_xi mode 2 with xi_ideal 1 and 0,
 * which returns each of them exactly.  Returns xi; dxi as oracle_calc_xi. */
double oracle_calc_xi_hess(const oracle_sys *s, const oracle_rx *rx, const double *c, double *dxi, double *h, int mode)
{
    const int na = s->natoms, n = 3 * na;
    oracle_rx r1 = *rx, r0 = *rx;
    r1.xi_ideal = 1.0;
    r0.xi_ideal = 0.0;
    double ds1[3 * ORACLE_MAX_ATOMS], ds0[3 * ORACLE_MAX_ATOMS];
    const double s1 = oracle_calc_xi(s, &r1, c, ds1, 2), s0 = oracle_calc_xi(s, &r0, c, ds0, 2);
    const double xi = oracle_calc_xi(s, rx, c, dxi, mode);
    static double h1[3 * ORACLE_MAX_ATOMS][3 * ORACLE_MAX_ATOMS], h0[3 * ORACLE_MAX_ATOMS][3 * ORACLE_MAX_ATOMS];
    memset(h1, 0, sizeof h1);
    memset(h0, 0, sizeof h0);
    /* d2 of a bond length: (|R|^2 delta_de - R_d R_e)/|R|^3, the sign of the
     * term it enters s1 with (forming -, breaking +), over the bond count */
    for (int pass = 0; pass < 2; ++pass) {
        const int nbond = pass == 0 ? rx->n_form : rx->n_break;
        const double sg = pass == 0 ? -1.0 : 1.0;
        for (int i = 0; i < nbond; ++i) {
            const int a = (pass == 0 ? rx->bond_form[i][0] : rx->bond_break[i][0]) - 1;
            const int b = (pass == 0 ? rx->bond_form[i][1] : rx->bond_break[i][1]) - 1;
            double R[3];
            for (int d = 0; d < 3; ++d) R[d] = c[3 * a + d] - c[3 * b + d];
            const double rinv = 1.0 / sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
            const double r3 = rinv * rinv * rinv;
            double blk[3][3];
            for (int d = 0; d < 3; ++d)
                for (int e = 0; e < 3; ++e) {
                    if (d == e) blk[d][e] = sg * (R[(d + 1) % 3] * R[(d + 1) % 3] + R[(d + 2) % 3] * R[(d + 2) % 3]) * r3;
                    else blk[d][e] = -sg * R[d < e ? d : e] * R[d < e ? e : d] * r3;
                    blk[d][e] = blk[d][e] / (double)(float)nbond;
                }
            for (int d = 0; d < 3; ++d)
                for (int e = 0; e < 3; ++e) {
                    h1[3 * a + d][3 * a + e] += blk[d][e];
                    h1[3 * a + d][3 * b + e] -= blk[d][e];
                    h1[3 * b + d][3 * a + e] -= blk[d][e];
                    h1[3 * b + d][3 * b + e] += blk[d][e];
                }
        }
    }
    /* reactant centre-of-mass distances: s0 = r_inf - |R|, mass-weighted */
    double com[4][3], mreac[4];
    for (int k = 0; k < rx->n_reac; ++k) {
        mreac[k] = 0.0;
        for (int i = 0; i < rx->reac_natoms[k]; ++i) mreac[k] += s->mass[rx->reac_atoms[k][i] - 1];
        com[k][0] = com[k][1] = com[k][2] = 0.0;
    }
    for (int k = 0; k < rx->n_reac; ++k)
        for (int i = 0; i < rx->reac_natoms[k]; ++i) {
            const int at = rx->reac_atoms[k][i] - 1;
            for (int j = 0; j < 3; ++j) com[k][j] = com[k][j] + s->mass[at] * c[3 * at + j] / mreac[k];
        }
    const int terms = (rx->n_reac * rx->n_reac - rx->n_reac) / 2;
    for (int i = 0; i < rx->n_reac; ++i)
        for (int j = i + 1; j < rx->n_reac; ++j) {
            double R[3];
            for (int d = 0; d < 3; ++d) R[d] = com[j][d] - com[i][d];
            const double rinv = 1.0 / sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
            const double r3 = rinv * rinv * rinv;
            double blk[3][3];
            for (int d = 0; d < 3; ++d)
                for (int e = 0; e < 3; ++e)
                    blk[d][e] = d == e ? -(R[(d + 1) % 3] * R[(d + 1) % 3] + R[(d + 2) % 3] * R[(d + 2) % 3]) * r3
                                       : R[d < e ? d : e] * R[d < e ? e : d] * r3;
            const int grp[2] = { i, j };
            for (int g1 = 0; g1 < 2; ++g1)
                for (int k = 0; k < rx->reac_natoms[grp[g1]]; ++k) {
                    const int at1 = rx->reac_atoms[grp[g1]][k] - 1;
                    for (int g2 = 0; g2 < 2; ++g2)
                        for (int l = 0; l < rx->reac_natoms[grp[g2]]; ++l) {
                            const int at2 = rx->reac_atoms[grp[g2]][l] - 1;
                            const double mf = s->mass[at1] / mreac[grp[g1]] * s->mass[at2] / mreac[grp[g2]];
                            const double sg = g1 == g2 ? 1.0 : -1.0;
                            for (int d = 0; d < 3; ++d)
                                for (int e = 0; e < 3; ++e)
                                    h0[3 * at1 + d][3 * at2 + e] += sg * (blk[d][e] * mf / (double)(float)terms);
                        }
                }
        }
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
            double v;
            if (mode == 1)
                v = ((s0 * h1[a][b] + ds0[b] * ds1[a] - ds1[b] * ds0[a] - s1 * h0[a][b]) * (s0 - s1) -
                     2.0 * (s0 * ds1[a] - s1 * ds0[a]) * (ds0[b] - ds1[b])) /
                    ((s0 - s1) * (s0 - s1) * (s0 - s1));
            else
                v = rx->xi_ideal * h1[a][b] + (1 - rx->xi_ideal) * h0[a][b];
            h[(size_t)a * n + b] = v;
        }
    return xi;
}

/* umbrella mode 0 (src/umbrella.f90:109-173) at the centroid c: xi (mode 1)
 * and its gradient into xi_real and dxi; the bias gradient
 * k_force (xi - xi_ideal) dxi, then the metric term dhams, added to every
 * bead of g (bead-major, as the reference adds them).  Returns the bias
 * energy 0.5 k_force (xi - xi_ideal)^2 - log(fs) / beta. */
double oracle_umbrella(const oracle_sys *s, const oracle_rx *rx, const double *c, double beta, double k_force,
                       double *dxi, double *g, double *xi_real)
{
    const int na = s->natoms, nb = s->nbeads, n = 3 * na;
    static double h[9 * ORACLE_MAX_ATOMS * ORACLE_MAX_ATOMS];
    const double xi = oracle_calc_xi_hess(s, rx, c, dxi, h, 1);
    if (xi_real) *xi_real = xi;
    const double delta = xi - rx->xi_ideal;
    double v_add = 0.5 * k_force * delta * delta;
    for (int k = 0; k < nb; ++k)
        for (int j = 0; j < na; ++j)
            for (int i = 0; i < 3; ++i) g[IDX(k, j, i, na)] = g[IDX(k, j, i, na)] + k_force * delta * dxi[3 * j + i];
    const double pi = 3.1415926535897932384;
    double fs2 = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < na; ++j) fs2 = fs2 + dxi[3 * j + i] * dxi[3 * j + i] / s->mass[j];
    const double coeff1 = 2.0 * pi * beta;
    fs2 = fs2 / coeff1;
    const double fs = sqrt(fs2);
    const double log_fs = log(fs);
    const double coeff2 = -1.0 / beta;
    v_add = v_add + coeff2 * log_fs;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < na; ++j) {
            double dhams = 0.0;
            for (int i2 = 0; i2 < 3; ++i2)
                for (int j2 = 0; j2 < na; ++j2)
                    dhams = dhams + h[(size_t)(3 * j2 + i2) * n + 3 * j + i] * dxi[3 * j2 + i2] / s->mass[j2];
            dhams = dhams * coeff2 / (coeff1 * fs2);
            for (int k = 0; k < nb; ++k) g[IDX(k, j, i, na)] = g[IDX(k, j, i, na)] + dhams;
        }
    return v_add;
}

/* Nose-Hoover chain thermostat, four thermostats (src/nhc.f90:35-169): the
 * kinetic energy in the centroid-momenta form sum p^2/(2m)/nbeads^2, then
 * nc = 5 substeps of the 3-stage Suzuki-Yoshida factorisation of the chain,
 * and all momenta scaled.  vnh[4]: thermostat velocities, qnh[4]: masses
 * (mdinit: nfree ekT Q^2 for the first, ekT Q^2 for the others, with
 * k_B = 0.316679e-5); the chain's own kT is the default-REAL quotient
 * 1.380649E-23/4.3597447E-18 times the temperature.  If eta is not NULL, the
 * thermostat positions are advanced as well (not part of the reference,
 * which does not need them; a test uses them for the conserved energy). */
void oracle_nhc(const oracle_sys *s, double *p, double *vnh, const double *qnh, double kelvin, double *eta)
{
    const int na = s->natoms, nb = s->nbeads, nfree = 3 * na;
    const double ekt = (double)(1.380649E-23f / 4.3597447E-18f) * kelvin;
    const int nc = 5, ns = 3;
    const double dtc = s->dt / (double)nc;
    double w[3];
    w[0] = 1.0 / (2.0 - pow(2.0, 1.0 / 3.0));
    w[1] = 1.0 - 2.0 * w[0];
    w[2] = w[0];
    double scale = 1.0, eksum = 0.0, gnh[4];
    for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j) {
            const double *pp = &p[IDX(j, i, 0, na)];
            eksum = eksum + (pp[0] * pp[0] + pp[1] * pp[1] + pp[2] * pp[2]) / (2.0 * s->mass[i]) / nb / nb;
        }
    for (int i = 0; i < nc; ++i)
        for (int j = 0; j < ns; ++j) {
            const double dts = w[j] * dtc, dt2 = 0.5 * dts, dt4 = 0.25 * dts, dt8 = 0.125 * dts;
            double e;
            gnh[3] = (qnh[2] * vnh[2] * vnh[2] - ekt) / qnh[3];
            vnh[3] = vnh[3] + gnh[3] * dt4;
            gnh[2] = (qnh[1] * vnh[1] * vnh[1] - ekt) / qnh[2];
            e = exp(-vnh[3] * dt8);
            vnh[2] = e * (vnh[2] * e + gnh[2] * dt4);
            gnh[1] = (qnh[0] * vnh[0] * vnh[0] - ekt) / qnh[1];
            e = exp(-vnh[2] * dt8);
            vnh[1] = e * (vnh[1] * e + gnh[1] * dt4);
            gnh[0] = (2.0 * eksum - (double)nfree * ekt) / qnh[0];
            e = exp(-vnh[1] * dt8);
            vnh[0] = e * (vnh[0] * e + gnh[0] * dt4);
            e = exp(-vnh[0] * dt2);
            scale = scale * e;
            eksum = eksum * e * e;
            if (eta)
                for (int k = 0; k < 4; ++k) eta[k] = eta[k] + vnh[k] * dt2;
            gnh[0] = (2.0 * eksum - (double)nfree * ekt) / qnh[0];
            e = exp(-vnh[1] * dt8);
            vnh[0] = e * (vnh[0] * e + gnh[0] * dt4);
            gnh[1] = (qnh[0] * vnh[0] * vnh[0] - ekt) / qnh[1];
            e = exp(-vnh[2] * dt8);
            vnh[1] = e * (vnh[1] * e + gnh[1] * dt4);
            gnh[2] = (qnh[1] * vnh[1] * vnh[1] - ekt) / qnh[2];
            e = exp(-vnh[3] * dt8);
            vnh[2] = e * (vnh[2] * e + gnh[2] * dt4);
            gnh[3] = (qnh[2] * vnh[2] * vnh[2] - ekt) / qnh[3];
            vnh[3] = vnh[3] + gnh[3] * dt4;
        }
    for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j)
            for (int k = 0; k < 3; ++k) p[IDX(j, i, k, na)] = scale * p[IDX(j, i, k, na)];
}

/* One verlet step with the Nose-Hoover chain (thermostat 2, src/verlet.f90:
 * 185-206 and :1086-1106): the chain, p -= dt/2 g, free ring polymer,
 * forces, p -= dt/2 g, the chain again. */
void oracle_verlet_nhc(const oracle_sys *s, double *q, double *p, double *g, double *vbead, double beta,
                       double *vnh, const double *qnh, double *eta)
{
    const size_t n = (size_t)3 * s->natoms * s->nbeads;
    const double kelvin = 1.0 / (beta * 0.316679e-5);
    oracle_nhc(s, p, vnh, qnh, kelvin, eta);
    for (size_t i = 0; i < n; ++i) p[i] = p[i] - 0.5 * s->dt * g[i];
    oracle_free_rp(s, q, p, beta);
    oracle_forces(s, q, vbead, g);
    for (size_t i = 0; i < n; ++i) p[i] = p[i] - 0.5 * s->dt * g[i];
    oracle_nhc(s, p, vnh, qnh, kelvin, eta);
}

/* One step of an umbrella-sampling window trajectory (verlet with
 * constrain = 0 in the sampling loops, src/calc_rate.f90:1569 and :1626):
 * p -= dt/2 g; free ring polymer; forces; umbrella mode 0 at the centroid of
 * the new positions, for the window rx->xi_ideal with k_force, adds its bias
 * to every bead's g (src/verlet.f90:1046); p -= dt/2 g; Andersen when istep
 * is a multiple of andersen_step, with the draw id 'draw'; transrot if
 * remove_com (the reference always applies it to windows).  Returns xi of
 * mode 1 at that centroid (the xi_real the sampling loop sums). */
double oracle_verlet_umbrella(const oracle_sys *s, const oracle_rx *rx, double *q, double *p, double *g,
                              double *vbead, double beta, double k_force, long istep, int andersen_step,
                              uint64_t seed, uint32_t traj, uint64_t draw, int remove_com)
{
    const int na = s->natoms, nb = s->nbeads;
    const size_t n = (size_t)3 * na * nb;
    for (size_t i = 0; i < n; ++i) p[i] = p[i] - 0.5 * s->dt * g[i];
    oracle_free_rp(s, q, p, beta);
    double c[3 * ORACLE_MAX_ATOMS], dxi[3 * ORACLE_MAX_ATOMS], xi = 0.0;
    oracle_centroid(s, q, c);
    oracle_forces(s, q, vbead, g);
    (void)oracle_umbrella(s, rx, c, beta, k_force, dxi, g, &xi);
    for (size_t i = 0; i < n; ++i) p[i] = p[i] - 0.5 * s->dt * g[i];
    if (andersen_step > 0 && istep % andersen_step == 0) oracle_andersen(s, p, beta, seed, traj, draw);
    /* verlet ends every step with constrain <= 0, windows included, with
     * transrot (src/verlet.f90:1299-1305) */
    if (remove_com) oracle_transrot(s, q, p);
    return xi;
}

/* mdinit with bias_mode 2 (src/mdinit.f90:79-98), after the caller's force
 * evaluation g: umbrella mode 0 at the centroid of q adds its bias gradient
 * to g and leaves the mode-1 gradient of xi in dxi, the constraint direction
 * of the parent's first step.  The momenta draw that follows in mdinit is the
 * caller's (oracle_andersen). */
void oracle_parent_begin(const oracle_sys *s, const oracle_rx *rx, const double *q, double *g, double *dxi,
                         double beta, double k_force)
{
    double c[3 * ORACLE_MAX_ATOMS];
    oracle_centroid(s, q, c);
    (void)oracle_umbrella(s, rx, c, beta, k_force, dxi, g, NULL);
}

/* One step of a recrossing parent (verlet with constrain = 1,
 * src/verlet.f90): p -= dt/2 g; free ring polymer; centroid c; constrain_q
 * with the constraint direction dxi of the previous step (:747); forces;
 * dxi <- gradient of xi (mode 2) at c, the centroid from before the
 * constraint (umbrella mode 1, :1048); p -= dt/2 g; constrain_p(dxi)
 * (:1081); Andersen every andersen_step steps (:1159-1165).  Returns the
 * constraint status (0 = converged), as the reference, which then gives the
 * step an energy penalty and goes on. */
int oracle_verlet_parent(const oracle_sys *s, const oracle_rx *rx, double *q, double *p, double *g, double *vbead,
                         double *dxi, double beta, long istep, int andersen_step, uint64_t seed, uint32_t traj)
{
    const int na = s->natoms, nb = s->nbeads;
    const size_t n = (size_t)3 * na * nb;
    for (size_t i = 0; i < n; ++i) p[i] = p[i] - 0.5 * s->dt * g[i];
    oracle_free_rp(s, q, p, beta);
    double c[3 * ORACLE_MAX_ATOMS];
    oracle_centroid(s, q, c);
    const int status = oracle_constrain_q(s, rx, c, dxi, q, p);
    oracle_forces(s, q, vbead, g);
    (void)oracle_calc_xi(s, rx, c, dxi, 2);
    for (size_t i = 0; i < n; ++i) p[i] = p[i] - 0.5 * s->dt * g[i];
    oracle_constrain_p(s, dxi, p);
    if (andersen_step > 0 && istep % andersen_step == 0) oracle_andersen(s, p, beta, seed, traj, (uint64_t)istep);
    return status;
}
