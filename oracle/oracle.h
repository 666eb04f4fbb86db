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

/* analytic surfaces (src/egrad_h3.f, src/egrad_ch4h.f) */
void oracle_h3_pote(const double r[3], double *pe, double dpe[3]);
void oracle_egrad_h3(const double *q, int natoms, int nbeads, double *v, double *dvdq);
void oracle_ch4h_pot(const double *x, double *energy, double *g);
void oracle_egrad_ch4h(const double *q, int natoms, int nbeads, double *v, double *dvdq);

#ifdef __cplusplus
}
#endif
#endif
