/*
 * oracle/oracle.h -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Plain-C restatement of the reference's RPMD hot path, used as the checker
 * for the HIP engine.  Only tests/, the smoke check and the CPU-baseline leg
 * of bench.py load liboracle.so; the product library never links it.
 * Layouts follow the reference: q(3,natoms,nbeads) column-major, i.e. the
 * coordinate index runs fastest, then the atom, then the bead.
 */
#ifndef CARACAL_ORACLE_H
#define CARACAL_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_MAX_ATOMS 16
#define ORACLE_MAX_BEADS 256
enum { ORACLE_SURF_H3 = 1, ORACLE_SURF_CH4H = 2, ORACLE_SURF_OH3 = 3, ORACLE_SURF_BRH2 = 4, ORACLE_SURF_O3 = 5 };

typedef struct {
    int surface, natoms, nbeads;
    double dt;
    double mass[ORACLE_MAX_ATOMS];
} oracle_sys;

/* BIMOLEC reaction coordinate definition, 1-based atom indices. */
typedef struct {
    int n_form, n_break;
    int bond_form[4][2], bond_break[4][2];
    double form_ref[4], break_ref[4];
    int n_reac;
    int reac_natoms[4];
    int reac_atoms[4][ORACLE_MAX_ATOMS];
    double r_inf, xi_ideal;
} oracle_rx;

/* QMDFF-form periodic force field (oracle/ff_box.c) */
#define ORACLE_FF_NPAR 10
typedef struct {
    int natoms, nbonds, nangles;
    const int *type;          /* 0 = O, 1 = H */
    const int *mol;           /* molecule index of each atom */
    const int *bonds;         /* 2 per bond, 0-based */
    const int *angles;        /* 3 per angle (i, vertex, k), 0-based */
    double box, rcut, zahn_a, zahn_par;
    double sr42[2][2], r0bj[2][2], zrep[2][2], arep[2][2];
} oracle_ff;
/* The reference's flexible water model (src/egrad_water.f90:36-333),
 * nwater molecules O,H,H; x, g: (3, 3*nwater).  Parameters in atomic units
 * (src/water_init.f90:83-106): r0_OH, r0_HH, De, a, k_HH, k_rtheta, k_rr,
 * q_H, q_O, sigma_OO, eps_OO. */
void oracle_water_params(double par[11]);
double oracle_water_eval(int nwater, const double *x, double *g, int periodic, double box, int zahn,
                         double zahn_a, double zahn_par, double coul_cut);

double oracle_ff_eval(const oracle_ff *f, const double *theta, const double *x, double *g, double *dth,
                      const double *v, double *dvth);

/* The reference's QMDFF (oracle/qmdff.c): the arrays prepare leaves in its
 * qmdff module, 1-based atom indices; element-pair tables [z_i-1][z_j-1]
 * (94 x 94); c6xy n x n (symmetric); rad in bohr (as prepare converts it). */
typedef struct {
    int n, nbond, nangl, ntors, nhb, nnci, nmols;
    const int *at;
    const double *molnum, *q;
    const int *bond;            /* 2 per bond */
    const double *vbond;        /* r0, k, exponent */
    const int *angl;            /* 3 per angle, vertex first */
    const double *vangl;        /* theta0, k */
    const int *tors;            /* 6 per torsion: i, j, k, l, terms, kind (2 = out of plane) */
    const double *vtors;        /* 14 per torsion */
    const int *hb;              /* 3 per listed hydrogen bond: A, B, H */
    const double *vhb;          /* 2 per listed hydrogen bond */
    const int *nci;             /* 3 per noncovalent pair: i, j, kind 1..6 */
    const double *c6xy, *r094, *sr42, *zab, *r0ab;
    double eps1[6], eps2[6];
    const double *rad, *scalehb;
    int periodic, zahn;
    double box[3], coul_cut, vdw_cut, cut_low, zahn_a, zahn_par;
} oracle_qmdff;
int oracle_qmdff_supported(const oracle_qmdff *f);
double oracle_qmdff_eval(const oracle_qmdff *f, const double *x, double *g, double *parts);

/* analytic surfaces (src/egrad_h3.f, src/egrad_ch4h.f) */
void oracle_h3_pote(const double r[3], double *pe, double dpe[3]);
void oracle_egrad_h3(const double *q, int natoms, int nbeads, double *v, double *dvdq);
void oracle_ch4h_pot(const double *x, double *energy, double *g);
void oracle_egrad_ch4h(const double *q, int natoms, int nbeads, double *v, double *dvdq);
/* src/egrad_oh3.f: atoms O, H1, H2, H3; r = the six distances in its order */
void oracle_oh3_pot(const double r[6], double *vtot, double dvdr[6]);
void oracle_egrad_oh3(const double *q, int natoms, int nbeads, double *v, double *dvdq);
/* src/egrad_brh2.f: atoms Ha, Br, Hc; r = (Br-Ha, Ha-Hc, Br-Hc), its R(1..3) */
void oracle_brh2_pot(const double r[3], double *energy, double dedr[3]);
void oracle_egrad_brh2(const double *q, int natoms, int nbeads, double *v, double *dvdq);
/* src/egrad_o3.f: three O atoms, bohr and hartree */
void oracle_egrad_o3(const double *q, int natoms, int nbeads, double *v, double *dvdq);

/* RPMD step and helpers (src/verlet.f90 and callees) */
void oracle_rfft(double *x, int n);
void oracle_forces(const oracle_sys *s, const double *q, double *v, double *g);
void oracle_free_rp(const oracle_sys *s, double *q, double *p, double beta);
void oracle_centroid(const oracle_sys *s, const double *q, double *c);
void oracle_transrot(const oracle_sys *s, const double *q, double *p);
double oracle_calc_xi(const oracle_sys *s, const oracle_rx *rx, const double *c, double *dxi, int mode);
double oracle_calc_xi_hess(const oracle_sys *s, const oracle_rx *rx, const double *c, double *dxi, double *h, int mode);
double oracle_umbrella(const oracle_sys *s, const oracle_rx *rx, const double *c, double beta, double k_force,
                       double *dxi, double *g, double *xi_real);
void oracle_parent_begin(const oracle_sys *s, const oracle_rx *rx, const double *q, double *g, double *dxi,
                         double beta, double k_force);
void oracle_nhc(const oracle_sys *s, double *p, double *vnh, const double *qnh, double kelvin, double *eta);
void oracle_verlet_nhc(const oracle_sys *s, double *q, double *p, double *g, double *vbead, double beta,
                       double *vnh, const double *qnh, double *eta);
double oracle_verlet_umbrella(const oracle_sys *s, const oracle_rx *rx, double *q, double *p, double *g,
                              double *vbead, double beta, double k_force, long istep, int andersen_step,
                              uint64_t seed, uint32_t traj, uint64_t draw, int remove_com);
int oracle_verlet_parent(const oracle_sys *s, const oracle_rx *rx, double *q, double *p, double *g, double *vbead,
                         double *dxi, double beta, long istep, int andersen_step, uint64_t seed, uint32_t traj);
int oracle_constrain_q(const oracle_sys *s, const oracle_rx *rx, const double *centroid, const double *dxi_act,
                       double *q, double *p);
void oracle_constrain_p(const oracle_sys *s, const double *dxi_act, double *p);
double oracle_child_weight(const oracle_sys *s, const oracle_rx *rx, const double *q, const double *p, double beta);
void oracle_andersen(const oracle_sys *s, double *p, double beta, uint64_t seed, uint32_t traj, uint64_t draw);
double oracle_verlet(const oracle_sys *s, const oracle_rx *rx, double *q, double *p, double *g, double *vbead,
                     double beta, int ensemble, int remove_com, long istep, int andersen_step, uint64_t seed,
                     uint32_t traj);

/* counter-based normals */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double oracle_andersen_normal(uint64_t seed, uint32_t traj, uint64_t draw, int coord, int bead);

#ifdef __cplusplus
}
#endif
/* rate theory after the umbrella samplings (rate.c) */
void oracle_umbrella_integration(int nwin, const double *xi_wins, const double *average, const double *variance,
                                 const double *k_force, double beta, double xi_min, double xi_max, int nbins,
                                 int umbr_traj, int umbr_step, double *dA, double *bin_coord, double *pmf);
void oracle_pmf_extrema(const double *pmf, int nbins, double xi_min, double xi_max, int minloc_zero, int *imax,
                        int *imin);
double oracle_rate_bimolecular(int npaths, double kappa, double r_inf, double beta, double m1, double m2,
                               double pmf_max, double pmf_min);

#endif
