// The reference's QMDFF as a force field of the box engine (qmdff.hip):
// device tables and the launch of energy and gradient for nbeads replicas.
#pragma once
#include "common.hpp"

namespace cam {

struct QmdffDev;

// Builds the device tables from a descriptor (checked by the caller).
int qmdff_build(const cam_qmdff_desc *d, QmdffDev **out);
void qmdff_free(QmdffDev *q);
int qmdff_natoms(const QmdffDev *q);
// Energies e[nbeads] and gradient g [bead][atom][3] (g may be NULL) of the
// positions x [bead][atom][3], on stream s.  T: the intermolecular pair
// arithmetic (double or float with compensated sums).
template <typename T>
int qmdff_launch(QmdffDev *q, const double *x, int nbeads, double *e, double *g, hipStream_t s);
// The reference's periodic wrap of every atom's ring polymer into the box
// after the free ring-polymer step (src/verlet.f90:540-640); no-op when the
// force field is not periodic.
int qmdff_wrap(QmdffDev *q, double *x, int nbeads, hipStream_t s);

}  // namespace cam
