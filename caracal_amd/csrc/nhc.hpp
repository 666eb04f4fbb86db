// Nose-Hoover chain thermostat shared by the trajectory engine and the box.
#pragma once

#include <hip/hip_runtime.h>

namespace cam {

// The Nose-Hoover chain of one system (a trajectory, or a box) over half a
// time step (src/nhc.f90:35-169): four thermostats with velocities vnh and masses qnh,
// nc = 5 substeps of the 3-stage Suzuki-Yoshida factorisation, on the
// centroid-momenta kinetic energy eksum.  Returns the factor that scales
// every momentum.  ekt is the chain's kT.
__device__ inline double nhc_chain(double *vnh, const double *qnh, double eksum, double ekt, int nfree, double dt)
{
    const double w0 = 1.0 / (2.0 - pow(2.0, 1.0 / 3.0));
    const double dtc = dt / 5.0;
    double scale = 1.0, gnh[4];
    for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double dts = (j == 1 ? 1.0 - 2.0 * w0 : w0) * dtc;
            const double dt2 = 0.5 * dts, dt4 = 0.25 * dts, dt8 = 0.125 * dts;
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
    return scale;
}

}  // namespace cam
