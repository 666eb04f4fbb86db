// Trajectory-batch RPMD engine: replaces the reference's per-step `verlet`
// (src/verlet.f90:65-1308) with its per-bead force loop, for a batch of
// independent trajectories that share a surface, an atom count and a bead
// count.
//
// Device layout of the state (q, p, dV/dq): [traj][3*natoms][nbeads] with
// the bead index fastest, so the lanes of one trajectory read consecutive
// addresses.  A block of 256 lanes holds 256/(L*nbeads) trajectories, where
// L is the number of lanes that share one bead (1 for H3, 4 for CH4H: see
// ch4h::egrad_quad).  One launch advances its trajectories by several steps
// with the state held in registers; the normal-mode step goes through LDS.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <type_traits>
#include <vector>

#include "philox.hpp"
#include "step_common.hpp"
#include "step_kernels.hpp"

namespace cam {
namespace {

// ---------------------------------------------------------------------------
// Per-trajectory helper kernels (one thread per trajectory).
// ---------------------------------------------------------------------------
struct TrajArgs {
    double *q, *p;
    const double *g, *vbead, *beta;
    int ntraj, natoms, nbeads;
    double dt;
    double mass[CAM_MAX_ATOMS];
    Rx rx;
};

// COM translation and rotation removal over all beads, as src/transrot.f90:36-237
// computes it (including its total mass of nbeads^2 * M).
__global__ void transrot_kernel(TrajArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ntraj) return;
    const int NA = a.natoms, NB = a.nbeads, NC = 3 * NA;
    double *q = a.q + (size_t)t * NC * NB, *p = a.p + (size_t)t * NC * NB;
    auto Q = [&](int d, int i, int k) -> double & { return q[(size_t)(3 * i + d) * NB + k]; };
    auto P = [&](int d, int i, int k) -> double & { return p[(size_t)(3 * i + d) * NB + k]; };
    double totmass = 0.0, vtot[3] = { 0, 0, 0 };
    for (int i = 0; i < NA; ++i) {
        const double w = a.mass[i];
        for (int k = 0; k < NB; ++k) {
            totmass = totmass + w;
            for (int j = 0; j < 3; ++j) vtot[j] = vtot[j] + (P(j, i, k) / a.mass[i]) * w;
        }
    }
    totmass = totmass * NB;
    for (int j = 0; j < 3; ++j) vtot[j] = vtot[j] / totmass;
    double xt = 0, yt = 0, zt = 0;
    for (int i = 0; i < NA; ++i) {
        const double w = a.mass[i];
        for (int k = 0; k < NB; ++k) {
            xt = xt + Q(0, i, k) * w;
            yt = yt + Q(1, i, k) * w;
            zt = zt + Q(2, i, k) * w;
        }
    }
    xt = xt / totmass;
    yt = yt / totmass;
    zt = zt / totmass;
    double mang[3] = { 0, 0, 0 };
    for (int i = 0; i < NA; ++i) {
        const double w = a.mass[i];
        for (int k = 0; k < NB; ++k) {
            const double vx = P(0, i, k) / a.mass[i], vy = P(1, i, k) / a.mass[i], vz = P(2, i, k) / a.mass[i];
            mang[0] = mang[0] + (Q(1, i, k) * vz - Q(2, i, k) * vy) * w;
            mang[1] = mang[1] + (Q(2, i, k) * vx - Q(0, i, k) * vz) * w;
            mang[2] = mang[2] + (Q(0, i, k) * vy - Q(1, i, k) * vx) * w;
        }
    }
    mang[0] = mang[0] - (yt * vtot[2] - zt * vtot[1]) * totmass;
    mang[1] = mang[1] - (zt * vtot[0] - xt * vtot[2]) * totmass;
    mang[2] = mang[2] - (xt * vtot[1] - yt * vtot[0]) * totmass;
    double xx = 0, xy = 0, xz = 0, yy = 0, yz = 0, zz = 0;
    for (int i = 0; i < NA; ++i) {
        const double w = a.mass[i];
        for (int k = 0; k < NB; ++k) {
            const double xd = Q(0, i, k) - xt, yd = Q(1, i, k) - yt, zd = Q(2, i, k) - zt;
            xx = xx + xd * xd * w;
            xy = xy + xd * yd * w;
            xz = xz + xd * zd * w;
            yy = yy + yd * yd * w;
            yz = yz + yd * zd * w;
            zz = zz + zd * zd * w;
        }
    }
    // tensor, column-major a(row,col) as the reference builds it
    double ten[3][3];
    ten[0][0] = yy + zz; ten[0][1] = -xy;    ten[0][2] = -xz;
    ten[1][0] = -xy;     ten[1][1] = xx + zz; ten[1][2] = -yz;
    ten[2][0] = -xz;     ten[2][1] = -yz;    ten[2][2] = xx + yy;
    if (NA <= 2) {
        ten[0][0] += 0.000001;
        ten[1][1] += 0.000001;
        ten[2][2] += 0.000001;
    }
    // Gauss-Jordan inversion with full pivoting (src/invert.f90)
    {
        int ipiv[3] = { 0, 0, 0 }, indxr[3], indxc[3];
        for (int i = 0; i < 3; ++i) {
            double big = 0.0;
            int irow = 0, icol = 0;
            for (int j = 0; j < 3; ++j)
                if (ipiv[j] != 1)
                    for (int kk = 0; kk < 3; ++kk)
                        if (ipiv[kk] == 0 && fabs(ten[j][kk]) >= big) {
                            big = fabs(ten[j][kk]);
                            irow = j;
                            icol = kk;
                        }
            ipiv[icol] += 1;
            if (irow != icol)
                for (int j = 0; j < 3; ++j) {
                    const double tmp = ten[irow][j];
                    ten[irow][j] = ten[icol][j];
                    ten[icol][j] = tmp;
                }
            indxr[i] = irow;
            indxc[i] = icol;
            const double piv = ten[icol][icol];
            ten[icol][icol] = 1.0;
            for (int j = 0; j < 3; ++j) ten[icol][j] = ten[icol][j] / piv;
            for (int j = 0; j < 3; ++j)
                if (j != icol) {
                    const double tmp = ten[j][icol];
                    ten[j][icol] = 0.0;
                    for (int kk = 0; kk < 3; ++kk) ten[j][kk] = ten[j][kk] - ten[icol][kk] * tmp;
                }
        }
        for (int i = 2; i >= 0; --i)
            if (indxr[i] != indxc[i])
                for (int kk = 0; kk < 3; ++kk) {
                    const double tmp = ten[kk][indxr[i]];
                    ten[kk][indxr[i]] = ten[kk][indxc[i]];
                    ten[kk][indxc[i]] = tmp;
                }
    }
    double vang[3];
    for (int i = 0; i < 3; ++i) {
        vang[i] = 0.0;
        for (int j = 0; j < 3; ++j) vang[i] = vang[i] + ten[i][j] * mang[j];
    }
    for (int i = 0; i < NA; ++i)
        for (int k = 0; k < NB; ++k) {
            double v[3];
            for (int j = 0; j < 3; ++j) v[j] = P(j, i, k) / a.mass[i] - vtot[j];
            const double xd = Q(0, i, k) - xt, yd = Q(1, i, k) - yt, zd = Q(2, i, k) - zt;
            v[0] = v[0] - vang[1] * zd + vang[2] * yd;
            v[1] = v[1] - vang[2] * xd + vang[0] * zd;
            v[2] = v[2] - vang[0] * yd + vang[1] * xd;
            for (int j = 0; j < 3; ++j) P(j, i, k) = v[j] * a.mass[i];
        }
}

// Centroid of trajectory t (3*natoms, atom-major).
__device__ void centroid_of(const TrajArgs &a, int t, double *c)
{
    const int NB = a.nbeads, NC = 3 * a.natoms;
    const double *q = a.q + (size_t)t * NC * NB;
    #pragma unroll
    for (int i = 0; i < NC; ++i) {
        double s = 0.0;
        for (int k = 0; k < NB; ++k) s = s + q[(size_t)i * NB + k];
        c[i] = s / NB;
    }
}

// constrain_q (src/constrain_q.f90:35-113) and optionally constrain_p
// (src/constrain_p.f90:36-76) on the centroid's dividing surface.
template <int NA>
__global__ void constrain_kernel(TrajArgs a, int project_p, int *status)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ntraj) return;
    const int NB = a.nbeads, NC = 3 * NA;
    double c[3 * NA], dxi_act[3 * NA], qt[3 * NA], dxn[3 * NA];
    centroid_of(a, t, c);
    (void)calc_xi<NA>(a.rx, c, dxi_act);
    double mult = 0.0, coeff = 0.0;
    int good = 0;
    for (int iter = 1; iter <= 200; ++iter) {
        coeff = mult * a.dt * a.dt / NB;
        #pragma unroll
        for (int i = 0; i < NC; ++i) qt[i] = c[i] + coeff * dxi_act[i] / a.mass[i / 3];
        const double sigma = calc_xi<NA>(a.rx, qt, dxn);
        double dsigma = 0.0;
        #pragma unroll
        for (int i = 0; i < NC; ++i) dsigma = dsigma + dxn[i] * a.dt * a.dt * dxi_act[i] / (a.mass[i / 3] * NB);
        const double dx = sigma / dsigma;
        mult = mult - dx;
        // single-precision literals 1.0E-8 and 1.0E-10 of the reference
        if (fabs(dx) < (double)1.0e-8f || fabs(sigma) < (double)1.0e-10f) break;
        if (iter == 200) {
            good = 1;
            break;
        }
    }
    if (status) status[t] = good;
    if (good) return;
    double *q = a.q + (size_t)t * NC * NB, *p = a.p + (size_t)t * NC * NB;
    // note: the shift uses the coefficient of the last iteration, computed
    // before its multiplier update, as the reference does
    #pragma unroll
    for (int i = 0; i < NC; ++i)
        for (int k = 0; k < NB; ++k) {
            q[(size_t)i * NB + k] = q[(size_t)i * NB + k] + coeff / a.mass[i / 3] * dxi_act[i];
            p[(size_t)i * NB + k] = p[(size_t)i * NB + k] + mult * a.dt / NB * dxi_act[i];
        }
    if (!project_p) return;
    double c1 = 0.0, c2 = 0.0;
    #pragma unroll
    for (int i = 0; i < NC; ++i)
        for (int k = 0; k < NB; ++k) c1 = c1 + dxi_act[i] * p[(size_t)i * NB + k] / a.mass[i / 3];
    #pragma unroll
    for (int i = 0; i < NC; ++i) c2 = c2 + dxi_act[i] * dxi_act[i] / a.mass[i / 3];
    const double lam = -c1 / c2 / NB;
    #pragma unroll
    for (int i = 0; i < NC; ++i)
        for (int k = 0; k < NB; ++k) p[(size_t)i * NB + k] = p[(size_t)i * NB + k] + lam * dxi_act[i];
}

// ---------------------------------------------------------------------------
// Recrossing parent (verlet with constrain = 1).  One wave per trajectory:
// its lanes cover the beads of each coordinate, and the scalar work of the
// constraint (xi, Newton's iteration) is repeated by every lane on the same
// values, so no lane waits for a broadcast.  pdxi[traj][3*natoms] carries the
// constraint direction from one step to the next; pfail holds per trajectory
// kPfConstraint (a constraint did not converge since the start), kPfCheck
// (the per-step trajectory check failed) and kPfThisStep (this step's
// constraint did not converge).
// ---------------------------------------------------------------------------

// Centroid of trajectory t in every lane: lane i < 3*natoms sums coordinate
// i over the beads in bead order (get_centroid, src/get_centroid.f90:66-85).
template <int NA>
__device__ void wave_centroid(const double *q, int nb, int lane, double *c)
{
    double own = 0.0;
    if (lane < 3 * NA) {
        for (int k = 0; k < nb; ++k) own = own + q[(size_t)lane * nb + k];
        own = own / nb;
    }
#pragma unroll
    for (int i = 0; i < 3 * NA; ++i) c[i] = __shfl(own, i, 64);
}

// rpmd_check's structure and energy conditions (src/rpmd_check.f90:74-98)
// for trajectory t, the same in every lane of its wave: a position that is
// NaN or above HUGE, or an energy (sum over the beads plus penalty) that is
// NaN, above HUGE or above epot_max.
__device__ bool trajectory_fails(const TrajArgs &a, int t, int lane, double penalty, double epot_max)
{
    const int NB = a.nbeads, NQ = 3 * a.natoms * NB;
    const double *q = a.q + (size_t)t * NQ;
    bool bad = false;
    for (int i = lane; i < NQ; i += 64) {
        const double v = q[i];
        bad = bad || v != v || v > DBL_MAX;
    }
    double ep = penalty;
    for (int k = 0; k < NB; ++k) ep = ep + a.vbead[(size_t)t * NB + k];
    bad = bad || ep != ep || ep > DBL_MAX || ep > epot_max;
    return __any(bad);
}

// mdinit with bias_mode 2 (src/mdinit.f90:79-98), after the forces: the
// umbrella mode-0 bias at the centroid is added to every bead's gradient,
// first its force-constant part, then its metric part, and the gradient of
// xi (mode 1) becomes the constraint direction of the first step.
template <int NA>
__global__ __launch_bounds__(64 * kParentWaves) void parent_begin_kernel(TrajArgs a, double *g, double k_force,
                                                                         double *pdxi, int *pfail)
{
    constexpr int NC = 3 * NA;
    const int lane = threadIdx.x % 64, t = blockIdx.x * kParentWaves + threadIdx.x / 64;
    if (t >= a.ntraj) return;
    const int NB = a.nbeads;
    double c[NC], dxi[NC], gk[NC], gh[NC];
    wave_centroid<NA>(a.q + (size_t)t * NC * NB, NB, lane, c);
    double xi1;
    (void)umbrella_bias<NA>(a.rx, c, a.beta[t], k_force, a.rx.xi_ideal, xi1, dxi, gk, gh);
    double *gt = g + (size_t)t * NC * NB;
#pragma unroll
    for (int i = 0; i < NC; ++i)
        for (int k = lane; k < NB; k += 64) gt[(size_t)i * NB + k] = (gt[(size_t)i * NB + k] + gk[i]) + gh[i];
    const double mine = pick_lane<NC>(dxi, lane);
    if (lane < NC) pdxi[(size_t)t * NC + lane] = mine;
    if (lane == 0) pfail[t] = 0;
}

// Between the free ring-polymer step and the forces (src/verlet.f90:649,
// :744-756): constrain_q (src/constrain_q.f90:35-113) on the centroid c of
// the propagated beads with the previous step's direction, then the next
// direction, the gradient of xi (mode 2) at the same c (umbrella mode 1,
// src/verlet.f90:1048, src/umbrella.f90:117-122).  A constraint that does not
// converge in 200 iterations leaves the beads as they are and is counted,
// where the reference prints a warning and gives the step an energy penalty.
template <int NA>
__global__ __launch_bounds__(64 * kParentWaves) void parent_constrain_kernel(TrajArgs a, double *pdxi, int *pfail)
{
    constexpr int NC = 3 * NA;
    const int lane = threadIdx.x % 64, t = blockIdx.x * kParentWaves + threadIdx.x / 64;
    if (t >= a.ntraj) return;
    const int NB = a.nbeads;
    double *q = a.q + (size_t)t * NC * NB, *p = a.p + (size_t)t * NC * NB;
    double c[NC], dprev[NC], qt[NC], dxn[NC];
    wave_centroid<NA>(q, NB, lane, c);
#pragma unroll
    for (int i = 0; i < NC; ++i) dprev[i] = pdxi[(size_t)t * NC + i];
    double mult = 0.0, coeff = 0.0;
    bool good = true;
    for (int iter = 1; iter <= 200; ++iter) {
        coeff = mult * a.dt * a.dt / NB;
#pragma unroll
        for (int i = 0; i < NC; ++i) qt[i] = c[i] + coeff * dprev[i] / a.mass[i / 3];
        const double sigma = calc_xi<NA>(a.rx, qt, dxn);
        double dsigma = 0.0;
        #pragma unroll
        for (int d = 0; d < 3; ++d)
            #pragma unroll
            for (int at = 0; at < NA; ++at)
                dsigma = dsigma + dxn[3 * at + d] * a.dt * a.dt * dprev[3 * at + d] / (a.mass[at] * NB);
        const double dx = sigma / dsigma;
        mult = mult - dx;
        // single-precision literals 1.0E-8 and 1.0E-10 of the reference
        if (fabs(dx) < (double)1.0e-8f || fabs(sigma) < (double)1.0e-10f) break;
        if (iter == 200) good = false;
    }
    if (good) {
        // the shift uses the coefficient of the last iteration, computed
        // before its multiplier update, as the reference does
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            const double dq = coeff / a.mass[i / 3] * dprev[i], dp = mult * a.dt / NB * dprev[i];
            for (int k = lane; k < NB; k += 64) {
                q[(size_t)i * NB + k] = q[(size_t)i * NB + k] + dq;
                p[(size_t)i * NB + k] = p[(size_t)i * NB + k] + dp;
            }
        }
    }
    if (lane == 0) pfail[t] = good ? (pfail[t] & ~kPfThisStep) : (pfail[t] | kPfConstraint | kPfThisStep);
    double dnew[NC];
    (void)calc_xi<NA>(a.rx, c, dnew);
    const double mine = pick_lane<NC>(dnew, lane);
    if (lane < NC) pdxi[(size_t)t * NC + lane] = mine;
}

// The end of a parent step (src/verlet.f90:1060-1083, :1159-1165): the
// second half kick, constrain_p (src/constrain_p.f90:36-76) with this step's
// direction, and the Andersen draw when the step number istep is a multiple
// of andersen_step.  istep counts from the start (cam_engine_begin_parent,
// first step 1), as the reference's verlet loops restart after each mdinit;
// the draw is keyed by the steps since the engine's creation, so restarts
// draw fresh momenta.
// Then rpmd_check as the recrossing driver calls it for a parent
// (src/rpmd_check.f90:74-98, traj_type 1 at src/recross_serial.f90:125-135
// and :290-299), from step check_after + 1 after the start on (never if
// check_after < 0): a position that is NaN or above HUGE, or an energy (the
// sum over the beads, plus verlet's 1e5 penalty after a failed constraint,
// src/verlet.f90:749-755) that is NaN, above HUGE or above epot_max, sets
// kPfCheck.  step_base[1] counts the steps since the start.
template <int NA>
__global__ __launch_bounds__(64 * kParentWaves) void parent_end_kernel(TrajArgs a, const double *pdxi, int *pfail,
                                                                       const uint64_t *step_base, int step_off,
                                                                       int andersen_step, uint64_t seed,
                                                                       double epot_max, int check_after)
{
    constexpr int NC = 3 * NA;
    const int lane = threadIdx.x % 64, t = blockIdx.x * kParentWaves + threadIdx.x / 64;
    if (t >= a.ntraj) return;
    const int NB = a.nbeads;
    double *p = a.p + (size_t)t * NC * NB;
    const double *g = a.g + (size_t)t * NC * NB;
    double dxi[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) dxi[i] = pdxi[(size_t)t * NC + i];
    double c1 = 0.0, c2 = 0.0;
#pragma unroll
    for (int i = 0; i < NC; ++i)
        for (int k = lane; k < NB; k += 64) {
            const double pn = p[(size_t)i * NB + k] - 0.5 * a.dt * g[(size_t)i * NB + k];
            p[(size_t)i * NB + k] = pn;
            c1 = c1 + dxi[i] * pn / a.mass[i / 3];
        }
    for (int off = 32; off > 0; off >>= 1) c1 += __shfl_xor(c1, off, 64);
    #pragma unroll
    for (int d = 0; d < 3; ++d)
        #pragma unroll
        for (int at = 0; at < NA; ++at) c2 = c2 + dxi[3 * at + d] * dxi[3 * at + d] / a.mass[at];
    const double lam = -c1 / c2 / NB;
    const uint64_t istep = step_base[1] + (uint64_t)step_off + 1;
    const uint64_t key = step_base[0] + (uint64_t)step_off + 1;
    const bool draw = andersen_step > 0 && istep % (uint64_t)andersen_step == 0;
    const double beta = a.beta[t];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const double sd = sqrt(a.mass[i / 3] / (beta / NB));
        for (int k = lane; k < NB; k += 64)
            p[(size_t)i * NB + k] = draw ? andersen_normal(seed, (uint32_t)t, key, i, k) * sd
                                         : p[(size_t)i * NB + k] + lam * dxi[i];
    }
    if (check_after >= 0 && step_base[1] + (uint64_t)step_off + 1 > (uint64_t)check_after) {
        const double penalty = (pfail[t] & kPfThisStep) ? 1.0e5 : 0.0;
        if (trajectory_fails(a, t, lane, penalty, epot_max) && lane == 0) pfail[t] = pfail[t] | kPfCheck;
    }
}

// ---------------------------------------------------------------------------
// Umbrella-sampling windows (verlet with constrain = 0 and the umbrella bias,
// src/calc_rate.f90:1556-1645), one window per trajectory: uwin[traj] =
// (xi_ideal, k_force) of its window, ustat[traj] = running sums of xi and xi^2
// (mode 1) since the start (src/calc_rate.f90:1639-1640).  Same wave layout
// as the parent kernels.
// ---------------------------------------------------------------------------

// Start of a window segment: with bias, the umbrella bias that mdinit adds
// to the fresh forces (src/mdinit.f90:90-96); without, the unbiased forces
// the sampling phase starts from (src/calc_rate.f90:1610-1618).  Clears the
// sums and the flags.
template <int NA>
__global__ __launch_bounds__(64 * kParentWaves) void umbrella_begin_kernel(TrajArgs a, double *g, const double *uwin,
                                                                           double *ustat, int *pfail, int bias)
{
    constexpr int NC = 3 * NA;
    const int lane = threadIdx.x % 64, t = blockIdx.x * kParentWaves + threadIdx.x / 64;
    if (t >= a.ntraj) return;
    const int NB = a.nbeads;
    if (bias) {
        double c[NC], dxi[NC], gk[NC], gh[NC], xi1;
        wave_centroid<NA>(a.q + (size_t)t * NC * NB, NB, lane, c);
        (void)umbrella_bias<NA>(a.rx, c, a.beta[t], uwin[2 * t + 1], uwin[2 * t], xi1, dxi, gk, gh);
        double *gt = g + (size_t)t * NC * NB;
#pragma unroll
        for (int i = 0; i < NC; ++i)
            for (int k = lane; k < NB; k += 64) gt[(size_t)i * NB + k] = (gt[(size_t)i * NB + k] + gk[i]) + gh[i];
    }
    if (lane == 0) {
        ustat[2 * t] = 0.0;
        ustat[2 * t + 1] = 0.0;
        pfail[t] = 0;
    }
}

// The end of a window step (src/verlet.f90:1040-1062, :1159-1165): the
// umbrella bias at the centroid of this step's positions added to every
// bead's gradient (umbrella mode 0, src/verlet.f90:1046), the running sums of
// xi, the second half kick, and the Andersen draw every andersen_step steps
// (counted from the start; draws keyed by the steps since creation).  Then the
// check of the umbrella loops (rpmd_check with traj_type 2 and 3,
// src/rpmd_check.f90:74-112): the parent's conditions, without the penalty,
// plus |xi - xi_ideal| > xi_tol.
template <int NA>
__global__ __launch_bounds__(64 * kParentWaves) void umbrella_end_kernel(TrajArgs a, double *g, const double *uwin,
                                                                         double *ustat, int *pfail,
                                                                         const uint64_t *step_base, int step_off,
                                                                         int andersen_step, uint64_t seed,
                                                                         double epot_max, double xi_tol, int check_after,
                                                                         unsigned *uhist, int hbins, double hmin,
                                                                         double hinv)
{
    constexpr int NC = 3 * NA;
    const int lane = threadIdx.x % 64, t = blockIdx.x * kParentWaves + threadIdx.x / 64;
    if (t >= a.ntraj) return;
    const int NB = a.nbeads;
    double c[NC], dxi[NC], gk[NC], gh[NC], xi1;
    wave_centroid<NA>(a.q + (size_t)t * NC * NB, NB, lane, c);
    const double xi_ideal = uwin[2 * t];
    (void)umbrella_bias<NA>(a.rx, c, a.beta[t], uwin[2 * t + 1], xi_ideal, xi1, dxi, gk, gh);
    const uint64_t istep = step_base[1] + (uint64_t)step_off + 1;
    const uint64_t key = step_base[0] + (uint64_t)step_off + 1;
    const bool draw = andersen_step > 0 && istep % (uint64_t)andersen_step == 0;
    const double beta = a.beta[t];
    double *gt = g + (size_t)t * NC * NB, *p = a.p + (size_t)t * NC * NB;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const double sd = sqrt(a.mass[i / 3] / (beta / NB));
        for (int k = lane; k < NB; k += 64) {
            const size_t o = (size_t)i * NB + k;
            const double gn = (gt[o] + gk[i]) + gh[i];
            gt[o] = gn;
            p[o] = draw ? andersen_normal(seed, (uint32_t)t, key, i, k) * sd : p[o] - 0.5 * a.dt * gn;
        }
    }
    if (lane == 0) {
        ustat[2 * t] = ustat[2 * t] + xi1;
        ustat[2 * t + 1] = ustat[2 * t + 1] + xi1 * xi1;
        count_xi(uhist, hbins, hmin, hinv, t, xi1);
    }
    if (check_after >= 0 && istep > (uint64_t)check_after) {
        const bool bad = trajectory_fails(a, t, lane, 0.0, epot_max) || fabs(xi1 - xi_ideal) > xi_tol;
        if (bad && lane == 0) pfail[t] = pfail[t] | kPfCheck;
    }
}

// Flux weight vs/fs of each child and the kappa denominator contribution
// (src/recross_serial.f90:189-233).
template <int NA>
__global__ void child_begin_kernel(TrajArgs a, double *weight, double *den_part)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ntraj) return;
    const int NB = a.nbeads, NC = 3 * NA;
    double c[3 * NA], dxi[3 * NA];
    centroid_of(a, t, c);
    (void)calc_xi<NA>(a.rx, c, dxi);
    const double *p = a.p + (size_t)t * NC * NB;
    double vs = 0.0;
    for (int i = 0; i < NC; ++i)
        for (int k = 0; k < NB; ++k) vs = vs + dxi[i] * p[(size_t)i * NB + k] / a.mass[i / 3];
    vs = vs / NB;
    double fs = 0.0;
    for (int i = 0; i < NC; ++i) fs = fs + dxi[i] * dxi[i] / a.mass[i / 3];
    const double beta = a.beta[t];
    fs = sqrt(fs / (2.0 * 3.141592653589793 * beta));
    weight[t] = vs / fs;
    den_part[t] = vs > 0.0 ? vs / fs : 0.0;
}

// num(t) += sum_traj weight * flag(t, traj), one block per step, fixed order.
// num[t] for t counted from the last begin_children: the run's first step is
// child step ctr[1] (the device counter, so that a replayed graph continues).
__global__ void kappa_reduce_kernel(const uint8_t *flag, const double *weight, int ntraj, double *num,
                                    const uint64_t *ctr)
{
    __shared__ double part[256];
    const int s = blockIdx.x;
    double acc = 0.0;
    for (int t = threadIdx.x; t < ntraj; t += 256) acc += flag[(size_t)s * ntraj + t] ? weight[t] : 0.0;
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) num[ctr[1] + (uint64_t)s] += part[0];
}

__global__ void sum_kernel(const double *v, int n, double *out)
{
    __shared__ double part[256];
    double acc = 0.0;
    for (int t = threadIdx.x; t < n; t += 256) acc += v[t];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out += part[0];
}

// ctr[0]: steps since creation (Andersen draws); ctr[1]: steps since begin_children.
__global__ void advance_kernel(uint64_t *ctr, int n)
{
    ctr[0] += (uint64_t)n;
    ctr[1] += (uint64_t)n;
}

__global__ void reset_child_kernel(uint64_t *ctr) { ctr[1] = 0; }

// Redraw momenta (andersen, src/andersen.f90:37-74).
__global__ void draw_kernel(double *p, const double *beta, int ntraj, int natoms, int nbeads, uint64_t seed,
                            uint64_t draw, int pair, TrajArgs a)
{
    const size_t NC = 3 * natoms;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)ntraj * NC * nbeads) return;
    const int k = (int)(i % nbeads);
    const int c = (int)((i / nbeads) % NC);
    const int t = (int)(i / (NC * nbeads));
    const int src = pair ? (t & ~1) : t;
    const double sd = sqrt(a.mass[c / 3] / (beta[src] / nbeads));
    const double z = andersen_normal(seed, (uint32_t)src, draw, c, k) * sd;
    p[i] = (pair && (t & 1)) ? -z : z;
}

// Observables per trajectory.
__global__ void observe_kernel(TrajArgs a, const int *pfail, double *epot, double *ekin, double *espring,
                               double *xi, int *flags, int with_xi)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ntraj) return;
    const int NB = a.nbeads, NC = 3 * a.natoms;
    const double *q = a.q + (size_t)t * NC * NB, *p = a.p + (size_t)t * NC * NB;
    double ep = 0.0, ek = 0.0, es = 0.0;
    int fl = 0;
    for (int k = 0; k < NB; ++k) ep += a.vbead[(size_t)t * NB + k];
    const double wn = NB / a.beta[t];
    for (int i = 0; i < NC; ++i) {
        const double m = a.mass[i / 3];
        for (int k = 0; k < NB; ++k) {
            const double pk = p[(size_t)i * NB + k], qk = q[(size_t)i * NB + k];
            ek += pk * pk / (2.0 * m);
            if (!isfinite(qk)) fl |= 1;
            if (NB > 1) {
                const double dq = qk - q[(size_t)i * NB + (k + 1) % NB];
                es += 0.5 * m * wn * wn * dq * dq;
            }
        }
    }
    if (pfail) fl |= (pfail[t] & 3) << 1;
    epot[t] = ep;
    ekin[t] = ek;
    espring[t] = es;
    flags[t] = fl;
    if (with_xi) {
        double c[3 * CAM_MAX_ATOMS];
        centroid_of(a, t, c);
        xi[t] = a.natoms == 3   ? calc_xi<3>(a.rx, c, nullptr)
                : a.natoms == 4 ? calc_xi<4>(a.rx, c, nullptr)
                                : calc_xi<6>(a.rx, c, nullptr);
    } else {
        xi[t] = 0.0;
    }
}

// Layout conversions between the ABI blocks (3,natoms,nbeads) and the device
// layout [3*natoms][nbeads].
__global__ void to_device_layout(const double *in, double *out, int ntraj, int nc, int nb)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)ntraj * nc * nb) return;
    const int c = (int)(i % nc);
    const int k = (int)((i / nc) % nb);
    const size_t t = i / ((size_t)nc * nb);
    out[(t * nc + c) * nb + k] = in[i];
}
__global__ void from_device_layout(const double *in, double *out, int ntraj, int nc, int nb)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)ntraj * nc * nb) return;
    const int c = (int)(i % nc);
    const int k = (int)((i / nc) % nb);
    const size_t t = i / ((size_t)nc * nb);
    out[i] = in[(t * nc + c) * nb + k];
}

}  // namespace
}  // namespace cam

// ---------------------------------------------------------------------------
// Host side of the engine.
// ---------------------------------------------------------------------------
using namespace cam;

// The kernels of one schedule: the fused step, the initial-force kernel, and
// the normal-mode step and force kernel of the split schedule.
struct Kernels {
    step_fn step = nullptr, nm = nullptr;
    force_fn force = nullptr, fsplit = nullptr;
    int step_tpb = 1;                 // trajectories per block of step
    bool folded = false;              // step is a folded kernel (takes a trajectory range)
};

struct cam_engine {
    cam_config cfg{};
    Rx rx{};
    hipStream_t stream = nullptr;
    hipStream_t slice_stream = nullptr;   // second trajectory slice of large folded launches
    hipEvent_t slice_fork = nullptr, slice_join = nullptr;
    double *q = nullptr, *p = nullptr, *g = nullptr, *vbead = nullptr, *beta = nullptr;
    double *tfwd = nullptr, *tinv = nullptr;
    double *tfold = nullptr;          // folded reference transform (nm_mode REFERENCE, nbeads >= 4)
    double *weight = nullptr, *den_part = nullptr, *den = nullptr, *num = nullptr;
    uint8_t *xi_flag = nullptr;
    double *cbuf = nullptr;           // CHILD: centroids of one fused launch
    double *pdxi = nullptr;           // PARENT: constraint directions [traj][3*natoms]
    int *pfail = nullptr;             // PARENT, UMBRELLA: trajectory flags (kPf*)
    double *uwin = nullptr;           // UMBRELLA: (xi_ideal, k_force) per trajectory
    double *ustat = nullptr;          // UMBRELLA: running sums of xi, xi^2 per trajectory
    double *nhc = nullptr;            // NHC: thermostat velocities [traj][4]
    long ustat_steps = 0;             // UMBRELLA: steps in those sums
    unsigned *uhist = nullptr;        // UMBRELLA: histograms of xi [traj][hbins] (cam_engine_set_umbrella_histogram)
    int hbins = 0;
    double hmin = 0.0, hinv = 0.0;
    bool parent_ready = false;        // PARENT, UMBRELLA: begun since the last set_state
    double pcheck_max = 0.0;          // PARENT, UMBRELLA: trajectory check
    double pcheck_xi_tol = 0.0;
    int pcheck_after = -1;
    uint64_t *step_base = nullptr;    // [0] steps since creation, [1] steps since begin_children
    double *scratch_io = nullptr;     // staging for layout conversion
    int num_cap = 0;                  // doubles in num (child steps)
    int flag_cap = 0;                 // steps of one run that xi_flag holds
    long child_steps = 0;             // host mirror of step_base[1]
    bool have_state = false;
    int spl = 1;                      // steps per launch
    int wpe = 1;
    int fold_nbuf = 1;                // LDS column buffers of a folded kernel (2 in a ragged batch)
    int nmk = CAM_NMK_VALU;           // normal-mode kernel of the current schedule
    Kernels kern;                     // kernels of the current schedule
    bool split = false;               // split schedule: force phase in its own kernel
    std::map<int, hipGraphExec_t> graphs;
    unsigned long graph_gen = 0;      // counts graph drops: graphs of a batch that hold this engine's launches
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    int last_launches = 0;            // step launches of the last cam_engine_run
    long last_steps = 0;              // steps of the last cam_engine_run
    int tpb = 1;                      // trajectories per block of the step kernel
};

namespace {
// Calls f with the engine's atom count as a compile-time constant (natoms is
// 3, 4 or 6 and matches the surface; checked at creation).
template <typename F>
void with_natoms(int natoms, F &&f)
{
    if (natoms == 3) f(std::integral_constant<int, 3>{});
    else if (natoms == 4) f(std::integral_constant<int, 4>{});
    else f(std::integral_constant<int, 6>{});
}
}  // namespace

namespace cam {
void nm_matrices(int nb, int mode, std::vector<double> &fwd, std::vector<double> &inv)
{
    // element (j,k) of the stored arrays is T[k][j]
    fwd.assign((size_t)nb * nb, 0.0);
    inv.assign((size_t)nb * nb, 0.0);
    const double pi = 3.141592653589793;
    if (mode == CAM_NM_REFERENCE) {
        // rfft/irfft: Re(DFT)/sqrt(N) -- the symmetric cosine matrix, used both ways
        const double f = std::sqrt(1.0 / nb);
        for (int kk = 0; kk < nb; ++kk)
            for (int j = 0; j < nb; ++j) {
                const long m = ((long)kk * j) % nb;
                const double c = std::cos(2.0 * pi * (double)m / nb) * f;
                fwd[(size_t)j * nb + kk] = c;
                inv[(size_t)j * nb + kk] = c;
            }
    } else {
        // orthogonal real normal-mode basis O[j][k] (bead j, mode k); mode k
        // and mode nb-k share the frequency 2 (nb/beta) sin(k pi / nb)
        std::vector<double> O((size_t)nb * nb);
        for (int j = 0; j < nb; ++j)
            for (int kk = 0; kk < nb; ++kk) {
                double v;
                if (kk == 0) v = std::sqrt(1.0 / nb);
                else if (2 * kk < nb) v = std::sqrt(2.0 / nb) * std::cos(2.0 * pi * j * kk / nb);
                else if (2 * kk == nb) v = ((j & 1) ? -1.0 : 1.0) / std::sqrt((double)nb);
                else v = std::sqrt(2.0 / nb) * std::sin(2.0 * pi * j * kk / nb);
                O[(size_t)j * nb + kk] = v;
            }
        // forward: y_k = sum_j O[j][k] x_j  -> T[k][j] = O[j][k]
        // inverse: x_k = sum_j O[k][j] y_j  -> Ti[k][j] = O[k][j]
        for (int kk = 0; kk < nb; ++kk)
            for (int j = 0; j < nb; ++j) {
                fwd[(size_t)j * nb + kk] = O[(size_t)j * nb + kk];
                inv[(size_t)j * nb + kk] = O[(size_t)kk * nb + j];
            }
    }
}

}  // namespace cam

namespace {

// The folded step kernel applies to the reference transform on NVE and child
// trajectories of the fused schedule without transrot (fold.hip).
bool fold_eligible(const cam_config *cfg)
{
    return cfg->nm_mode == CAM_NM_REFERENCE && cfg->nbeads >= 4 &&
           (cfg->ensemble == CAM_ENS_NVE || cfg->ensemble == CAM_ENS_CHILD) && !cfg->remove_com &&
           cfg->schedule != CAM_SCHED_SPLIT;
}

// ... and has a kernel for the surface, bead count and precision.
bool fold_available(const cam_config *cfg)
{
    int tpb = 0;
    return fold_eligible(cfg) &&
           fold_step_kernel_for(cfg->surface, cfg->precision, cfg->nbeads, cfg->ensemble, 1, 1, &tpb) != nullptr;
}

bool select_kernels(cam_engine *e, int wpe, int nmk, Kernels &k)
{
    const cam_config *cfg = &e->cfg;
    const int nb = cfg->nbeads;
    // a parent's or window's split-schedule normal-mode kernel is the NVE
    // one: its constraint or bias and the end of its step are kernels of
    // their own.  Windows also have a fused step; parents do not.
    const bool steered = cfg->ensemble == CAM_ENS_PARENT || cfg->ensemble == CAM_ENS_UMBRELLA;
    const int ens_nm = steered || cfg->ensemble == CAM_ENS_NHC ? CAM_ENS_NVE : cfg->ensemble;
    const int ens = cfg->ensemble == CAM_ENS_PARENT ? CAM_ENS_NVE : cfg->ensemble;
    const bool mfma = nmk == CAM_NMK_MFMA;
    bool ok;
    if (cfg->surface == CAM_SURFACE_H3)
        ok = pick_kernels_h3(cfg->precision, nb, ens, ens_nm, wpe, mfma, k.step, k.force, k.nm, k.fsplit);
    else if (cfg->surface == CAM_SURFACE_CH4H)
        ok = pick_kernels_ch4h(cfg->precision, nb, ens, ens_nm, wpe, mfma, k.step, k.force, k.nm, k.fsplit);
    else
        ok = pick_kernels_fp64(cfg->surface, nb, ens, ens_nm, wpe, mfma, k.step, k.force, k.nm, k.fsplit);
    k.step_tpb = e->tpb;
    if (ok && nmk == CAM_NMK_FOLDED) {
        int tpb = 0;
        const step_fn f =
            fold_eligible(cfg) ? fold_step_kernel_for(cfg->surface, cfg->precision, nb, ens, wpe, e->fold_nbuf, &tpb)
                               : nullptr;
        if (!f) return false;
        k.step = f;
        k.step_tpb = tpb;
        k.folded = true;
    }
    return ok;
}

TrajArgs traj_args(cam_engine *e)
{
    TrajArgs a{};
    a.q = e->q;
    a.p = e->p;
    a.g = e->g;
    a.vbead = e->vbead;
    a.beta = e->beta;
    a.ntraj = e->cfg.ntraj;
    a.natoms = e->cfg.natoms;
    a.nbeads = e->cfg.nbeads;
    a.dt = e->cfg.dt;
    std::memcpy(a.mass, e->cfg.mass, sizeof a.mass);
    a.rx = e->rx;
    return a;
}

int run_forces(cam_engine *e)
{
    const int grid = (e->cfg.ntraj + e->tpb - 1) / e->tpb;
    hipLaunchKernelGGL(e->kern.force, dim3(grid), dim3(kBlock), 0, e->stream, e->q, e->g, e->vbead, e->cfg.ntraj);
    CAM_LAUNCH_CHECK();
    return CAM_OK;
}

// Destroys the cached graphs.  A replay of one may still be in flight on the
// engine's stream (cam_engine_run is asynchronous), so the stream is drained
// first: no executable graph is destroyed while the GPU runs it.
void drop_graphs(cam_engine *e)
{
    ++e->graph_gen;
    if (e->graphs.empty()) return;
    (void)hipStreamSynchronize(e->stream);
    for (auto &kv : e->graphs) (void)hipGraphExecDestroy(kv.second);
    e->graphs.clear();
}

// Room for the flags of an nsteps run and for the numerator of child steps
// up to 'total'; the numerator keeps its contents when it grows.  Captured
// graphs refer to the buffers, so growth drops them.
int ensure_child_buffers(cam_engine *e, int nsteps, long total)
{
    if (nsteps > e->flag_cap) {
        // the graphs that write the old buffer go first (drop_graphs waits
        // for the stream, so nothing still uses it when it is freed)
        drop_graphs(e);
        CAM_HIP(hipStreamSynchronize(e->stream));
        if (e->xi_flag) CAM_HIP(hipFree(e->xi_flag));
        e->xi_flag = nullptr;
        CAM_HIP(hipMalloc(&e->xi_flag, (size_t)nsteps * e->cfg.ntraj));
        e->flag_cap = nsteps;
    }
    if (total > e->num_cap) {
        const long cap = std::max<long>(total, std::max<long>(2L * e->num_cap, 8192L));
        CAM_CHECK_ARG(cap <= (1L << 30), "too many child steps (%ld)", total);
        drop_graphs(e);
        double *nn = nullptr;
        CAM_HIP(hipMalloc(&nn, (size_t)cap * sizeof(double)));
        CAM_HIP(hipMemsetAsync(nn, 0, (size_t)cap * sizeof(double), e->stream));
        if (e->num) {
            CAM_HIP(hipMemcpyAsync(nn, e->num, (size_t)e->num_cap * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
            CAM_HIP(hipStreamSynchronize(e->stream));
            CAM_HIP(hipFree(e->num));
        }
        e->num = nn;
        e->num_cap = (int)cap;
    }
    return CAM_OK;
}

// Enqueue one run of nsteps (used both for capture and eagerly).
StepArgs step_args(cam_engine *e)
{
    StepArgs a{};
    a.q = e->q;
    a.p = e->p;
    a.g = e->g;
    a.vbead = e->vbead;
    a.beta = e->beta;
    a.tfwd = e->tfwd;
    a.tinv = e->tinv;
    a.tfold = e->tfold;
    a.xi_flag = e->xi_flag;
    a.cbuf = e->cbuf;
    a.pdxi = e->pdxi;
    a.pfail = e->pfail;
    a.ustat = e->ustat;
    a.uwin = e->uwin;
    a.nhc = e->nhc;
    a.uhist = e->uhist;
    a.hbins = e->hbins;
    a.hmin = e->hmin;
    a.hinv = e->hinv;
    a.nose_q = e->cfg.nose_q;
    a.epot_max = e->pcheck_max;
    a.xi_tol = e->pcheck_xi_tol;
    a.check_after = e->pcheck_after;
    a.step_base = e->step_base;
    a.ntraj = e->cfg.ntraj;
    a.tbegin = 0;
    a.tend = e->cfg.ntraj;
    // ANDERSEN draws every andersen_step >= 1 steps; a window's 0 means never
    // (the fused window step tests andersen_step > 0)
    a.andersen_step = e->cfg.ensemble == CAM_ENS_UMBRELLA ? e->cfg.andersen_step
                                                          : (e->cfg.andersen_step > 0 ? e->cfg.andersen_step : 1);
    a.seed = e->cfg.seed;
    a.dt = e->cfg.dt;
    std::memcpy(a.mass, e->cfg.mass, sizeof a.mass);
    a.rx = e->rx;
    return a;
}

// The kernel launches of nsteps steps on the state a points to: fused (spl
// steps per launch, transrot after each launch when requested) or split (per
// step the normal-mode kernel, then the force kernel; one last half kick).
hipError_t launch_steps(cam_engine *e, const Kernels &kn, bool split, int spl, StepArgs a, int nsteps,
                        const TrajArgs *transrot, bool slices = true)
{
    const int grid = (e->cfg.ntraj + e->tpb - 1) / e->tpb;
    hipError_t err = hipSuccess;
    if (e->cfg.ensemble == CAM_ENS_PARENT) {
        // per step: kick and free ring polymer; constraint; forces; second
        // half kick, constrain_p and Andersen
        TrajArgs ta = traj_args(e);
        ta.q = a.q;
        ta.p = a.p;
        ta.g = a.g;
        const int pgrid = (e->cfg.ntraj + kParentWaves - 1) / kParentWaves;
        for (int j = 0; j < nsteps && err == hipSuccess; ++j) {
            a.nsteps = 1;
            a.step_off = 0;                // no previous step to finish inside the kernel
            hipLaunchKernelGGL(kn.nm, dim3(grid), dim3(kBlock), 0, e->stream, a);
            err = hipGetLastError();
            if (err != hipSuccess) break;
            with_natoms(e->cfg.natoms, [&](auto na) {
                hipLaunchKernelGGL(parent_constrain_kernel<decltype(na)::value>, dim3(pgrid), dim3(64 * kParentWaves), 0, e->stream, ta, a.pdxi, a.pfail);
            });
            err = hipGetLastError();
            if (err != hipSuccess) break;
            hipLaunchKernelGGL(kn.fsplit, dim3(grid), dim3(kBlock), 0, e->stream, a.q, a.g, a.vbead, a.ntraj);
            err = hipGetLastError();
            if (err != hipSuccess) break;
            const int as = e->cfg.andersen_step;
            with_natoms(e->cfg.natoms, [&](auto na) {
                hipLaunchKernelGGL(parent_end_kernel<decltype(na)::value>, dim3(pgrid), dim3(64 * kParentWaves), 0, e->stream, ta, a.pdxi, a.pfail, a.step_base, j, as, a.seed, e->pcheck_max, e->pcheck_after);
            });
            err = hipGetLastError();
        }
        return err;
    }
    if (e->cfg.ensemble == CAM_ENS_UMBRELLA && split) {
        // per step: kick and free ring polymer; forces; bias, sums, second half kick, Andersen
        TrajArgs ta = traj_args(e);
        ta.q = a.q;
        ta.p = a.p;
        ta.g = a.g;
        const int pgrid = (e->cfg.ntraj + kParentWaves - 1) / kParentWaves;
        const int as = e->cfg.andersen_step;
        for (int j = 0; j < nsteps && err == hipSuccess; ++j) {
            a.nsteps = 1;
            a.step_off = 0;
            hipLaunchKernelGGL(kn.nm, dim3(grid), dim3(kBlock), 0, e->stream, a);
            err = hipGetLastError();
            if (err != hipSuccess) break;
            hipLaunchKernelGGL(kn.fsplit, dim3(grid), dim3(kBlock), 0, e->stream, a.q, a.g, a.vbead, a.ntraj);
            err = hipGetLastError();
            if (err != hipSuccess) break;
            with_natoms(e->cfg.natoms, [&](auto na) {
                hipLaunchKernelGGL(umbrella_end_kernel<decltype(na)::value>, dim3(pgrid), dim3(64 * kParentWaves), 0, e->stream, ta, a.g, e->uwin, a.ustat, a.pfail, a.step_base, j, as, a.seed, e->pcheck_max, e->pcheck_xi_tol, e->pcheck_after, a.uhist, a.hbins, a.hmin, a.hinv);
            });
            err = hipGetLastError();
            if (err == hipSuccess && transrot) {
                TrajArgs tr = *transrot;
                tr.q = a.q;
                tr.p = a.p;
                hipLaunchKernelGGL(transrot_kernel, dim3((e->cfg.ntraj + 127) / 128), dim3(128), 0, e->stream, tr);
                err = hipGetLastError();
            }
        }
        return err;
    }
    if (split) {
        for (int j = 0; j < nsteps && err == hipSuccess; ++j) {
            a.nsteps = 1;
            a.step_off = j;
            hipLaunchKernelGGL(kn.nm, dim3(grid), dim3(kBlock), 0, e->stream, a);
            err = hipGetLastError();
            if (err != hipSuccess) break;
            hipLaunchKernelGGL(kn.fsplit, dim3(grid), dim3(kBlock), 0, e->stream, a.q, a.g, a.vbead, a.ntraj);
            err = hipGetLastError();
        }
        if (err != hipSuccess) return err;
        a.nsteps = nsteps;
        a.andersen_step = e->cfg.ensemble == CAM_ENS_ANDERSEN ? e->cfg.andersen_step : 0;
        const size_t n = (size_t)e->cfg.ntraj * 3 * e->cfg.natoms * e->cfg.nbeads;
        hipLaunchKernelGGL(final_kick_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, e->stream, a,
                           e->cfg.natoms, e->cfg.nbeads);
        return hipGetLastError();
    }
    const int tgrid = (e->cfg.ntraj + 127) / 128;
    const int sgrid = (e->cfg.ntraj + kn.step_tpb - 1) / kn.step_tpb;
    // The launches of one trajectory slice [a.tbegin, a.tend) on stream st.
    const auto slice = [&](StepArgs b, hipStream_t st) {
        const int grid = (b.tend - b.tbegin + kn.step_tpb - 1) / kn.step_tpb;
        hipError_t rc = hipSuccess;
        for (int done = 0; done < nsteps && rc == hipSuccess; done += b.nsteps) {
            b.nsteps = std::min(spl, nsteps - done);
            b.step_off = done;
            hipLaunchKernelGGL(kn.step, dim3(grid), dim3(kBlock), 0, st, b);
            rc = hipGetLastError();
            if (rc == hipSuccess && e->cfg.ensemble == CAM_ENS_CHILD) {
                const unsigned xgrid = (unsigned)(((size_t)b.nsteps * (b.tend - b.tbegin) + 255) / 256);
                with_natoms(e->cfg.natoms, [&](auto na) {
                    hipLaunchKernelGGL(xi_flag_kernel<decltype(na)::value>, dim3(xgrid), dim3(256), 0, st, b);
                });
                rc = hipGetLastError();
            }
            if (rc == hipSuccess && transrot) {
                hipLaunchKernelGGL(transrot_kernel, dim3(tgrid), dim3(128), 0, st, *transrot);
                rc = hipGetLastError();
            }
        }
        return rc;
    };
    // A folded launch runs every block for all its steps, so the blocks of
    // its last round (when the grid is not a whole number of rounds) leave
    // part of the chip idle until they end.  With at least two blocks per CU
    // the trajectories are issued as two slices on two streams: the slices
    // are independent, and one slice's blocks fill the other's last round.
    // Not when the caller's stream is itself a branch forked from another
    // stream (a ragged batch's bucket, whose neighbours fill the chip): the
    // slices fork only from a stream that starts its own work.
    if (!slices || !kn.folded || transrot || sgrid < 2 * kSliceMinBlocks || !e->slice_stream)
        return slice(a, e->stream);
    StepArgs lo = a, hi = a;
    lo.tend = hi.tbegin = (sgrid / 2) * kn.step_tpb;
    err = hipEventRecord(e->slice_fork, e->stream);
    if (err == hipSuccess) err = hipStreamWaitEvent(e->slice_stream, e->slice_fork, 0);
    if (err == hipSuccess) err = slice(lo, e->stream);
    if (err == hipSuccess) err = slice(hi, e->slice_stream);
    // joined in every case, so that a capture is never left forked
    hipError_t jn = hipEventRecord(e->slice_join, e->slice_stream);
    if (jn == hipSuccess) jn = hipStreamWaitEvent(e->stream, e->slice_join, 0);
    return err != hipSuccess ? err : jn;
}

// slices: the folded launches may fork into two trajectory slices
// (launch_steps); false for a stream that is itself a forked branch.
int enqueue_run(cam_engine *e, int nsteps, bool slices)
{
    const TrajArgs ta = traj_args(e);
    CAM_HIP(launch_steps(e, e->kern, e->split, e->spl, step_args(e), nsteps, e->cfg.remove_com ? &ta : nullptr,
                         slices));
    if (e->cfg.ensemble == CAM_ENS_CHILD) {
        hipLaunchKernelGGL(kappa_reduce_kernel, dim3(nsteps), dim3(256), 0, e->stream, e->xi_flag, e->weight,
                           e->cfg.ntraj, e->num, e->step_base);
        CAM_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(1), 0, e->stream, e->step_base, nsteps);
    CAM_LAUNCH_CHECK();
    return CAM_OK;
}

}  // namespace

extern "C" {

int cam_engine_create(const cam_config *cfg, cam_engine **out)
{
    CAM_CHECK_ARG(cfg && out, "null argument");
    *out = nullptr;
    const int surf = cfg->surface;
    CAM_CHECK_ARG(surf == CAM_SURFACE_H3 || surf == CAM_SURFACE_CH4H || surf == CAM_SURFACE_OH3 ||
                      surf == CAM_SURFACE_BRH2 || surf == CAM_SURFACE_O3,
                  "unknown surface %d", surf);
    const int na = surf == CAM_SURFACE_OH3 ? 4 : surf == CAM_SURFACE_CH4H ? 6 : 3;
    CAM_CHECK_ARG(cfg->natoms == na, "surface %d needs natoms = %d (got %d)", surf, na, cfg->natoms);
    const int nb = cfg->nbeads;
    CAM_CHECK_ARG(bead_count_supported(nb), "nbeads must be 1, 2, 4, 6, 8, 12, 16, 24, 32, 48 or 64 (got %d)", nb);
    CAM_CHECK_ARG(cfg->ntraj >= 1, "ntraj must be >= 1 (got %d)", cfg->ntraj);
    CAM_CHECK_ARG(cfg->precision == CAM_PREC_FP64 || cfg->precision == CAM_PREC_MIXED, "bad precision");
    CAM_CHECK_ARG(surf == CAM_SURFACE_H3 || surf == CAM_SURFACE_CH4H || cfg->precision == CAM_PREC_FP64,
                  "the OH3, BrH2 and O3 surfaces run in FP64 only");
    CAM_CHECK_ARG(cfg->nm_mode == CAM_NM_REFERENCE || cfg->nm_mode == CAM_NM_ORTHOGONAL, "bad nm_mode");
    CAM_CHECK_ARG(cfg->ensemble >= CAM_ENS_NVE && cfg->ensemble <= CAM_ENS_NHC, "bad ensemble");
    CAM_CHECK_ARG(cfg->ensemble != CAM_ENS_NHC || cfg->nose_q > 0.0, "ensemble NHC needs nose_q > 0");
    CAM_CHECK_ARG(cfg->ensemble != CAM_ENS_NHC ||
                      (cfg->schedule != CAM_SCHED_SPLIT && (cfg->nm_kernel == CAM_NMK_AUTO || cfg->nm_kernel == CAM_NMK_VALU)),
                  "ensemble NHC runs on the fused schedule with the vector-ALU transform");
    const bool steered = cfg->ensemble == CAM_ENS_PARENT || cfg->ensemble == CAM_ENS_UMBRELLA;
    CAM_CHECK_ARG(cfg->nm_kernel >= CAM_NMK_AUTO && cfg->nm_kernel <= CAM_NMK_FOLDED, "bad nm_kernel");
    CAM_CHECK_ARG(cfg->nm_kernel != CAM_NMK_MFMA || mfma_bead_count(nb),
                  "nm_kernel MFMA needs nbeads 16, 32 or 64 (got %d)", nb);
    CAM_CHECK_ARG(cfg->nm_kernel != CAM_NMK_FOLDED || fold_eligible(cfg),
                  "nm_kernel FOLDED needs nm_mode REFERENCE, nbeads >= 4, ensemble NVE or CHILD, no remove_com "
                  "and the fused schedule");
    CAM_CHECK_ARG(cfg->nm_kernel != CAM_NMK_FOLDED || fold_available(cfg),
                  "nm_kernel FOLDED has no kernel for this surface, bead count and precision");
    CAM_CHECK_ARG(cfg->waves_per_eu >= 0 && cfg->waves_per_eu <= 2, "waves_per_eu must be 0, 1 or 2");
    CAM_CHECK_ARG(cfg->schedule >= CAM_SCHED_AUTO && cfg->schedule <= CAM_SCHED_SPLIT, "bad schedule");
    CAM_CHECK_ARG(cfg->schedule != CAM_SCHED_SPLIT || !cfg->remove_com || cfg->ensemble == CAM_ENS_UMBRELLA,
                  "remove_com needs the fused schedule (transrot after every step)");
    CAM_CHECK_ARG(cfg->reserved == 0, "reserved must be 0");
    CAM_CHECK_ARG(cfg->dt > 0.0, "dt must be > 0");
    CAM_CHECK_ARG(cfg->ensemble != CAM_ENS_ANDERSEN || cfg->andersen_step >= 1, "andersen_step must be >= 1");
    CAM_CHECK_ARG(!steered || cfg->andersen_step >= 0, "andersen_step must be >= 0");
    CAM_CHECK_ARG(cfg->ensemble != CAM_ENS_PARENT || cfg->schedule != CAM_SCHED_FUSED,
                  "ensemble PARENT runs on the split schedule (its constraint sits between propagation and forces)");
    for (int i = 0; i < na; ++i) CAM_CHECK_ARG(cfg->mass[i] > 0.0, "mass[%d] must be > 0", i);
    CAM_CHECK_ARG(!(cfg->remove_com && (cfg->ensemble == CAM_ENS_CHILD || cfg->ensemble == CAM_ENS_PARENT)),
                  "remove_com applies to trajectories with verlet's constrain <= 0 only (dynamic.x, windows)");
    const cam_reaction &rr = cfg->reaction;
    Rx rx{};
    if (cfg->ensemble == CAM_ENS_CHILD || steered) {
        CAM_CHECK_ARG(rr.n_form >= 0 && rr.n_form <= CAM_MAX_BONDS && rr.n_break >= 0 && rr.n_break <= CAM_MAX_BONDS &&
                          rr.n_form + rr.n_break > 0,
                      "reaction needs 1..%d forming and breaking bonds", CAM_MAX_BONDS);
        CAM_CHECK_ARG(rr.n_reac >= 2 && rr.n_reac <= CAM_MAX_REAC, "reaction needs 2..%d reactants", CAM_MAX_REAC);
    }
    rx.n_form = rr.n_form;
    rx.n_break = rr.n_break;
    rx.n_reac = rr.n_reac;
    rx.r_inf = rr.r_inf;
    rx.xi_ideal = rr.xi_ideal;
    for (int i = 0; i < rr.n_form && i < CAM_MAX_BONDS; ++i) {
        for (int j = 0; j < 2; ++j) {
            CAM_CHECK_ARG(rr.bond_form[i][j] >= 1 && rr.bond_form[i][j] <= na, "bond_form atom out of range");
            rx.form[i][j] = rr.bond_form[i][j] - 1;
        }
        rx.form_ref[i] = rr.form_ref[i];
    }
    for (int i = 0; i < rr.n_break && i < CAM_MAX_BONDS; ++i) {
        for (int j = 0; j < 2; ++j) {
            CAM_CHECK_ARG(rr.bond_break[i][j] >= 1 && rr.bond_break[i][j] <= na, "bond_break atom out of range");
            rx.brk[i][j] = rr.bond_break[i][j] - 1;
        }
        rx.break_ref[i] = rr.break_ref[i];
    }
    for (int k = 0; k < rr.n_reac && k < CAM_MAX_REAC; ++k) {
        CAM_CHECK_ARG(rr.reac_natoms[k] >= 1 && rr.reac_natoms[k] <= na, "reactant size out of range");
        rx.reac_n[k] = rr.reac_natoms[k];
        rx.reac_mass[k] = 0.0;
        for (int i = 0; i < rr.reac_natoms[k]; ++i) {
            CAM_CHECK_ARG(rr.reac_atoms[k][i] >= 1 && rr.reac_atoms[k][i] <= na, "reactant atom out of range");
            rx.reac_at[k][i] = rr.reac_atoms[k][i] - 1;
            rx.reac_mass[k] += cfg->mass[rx.reac_at[k][i]];
        }
    }
    for (int k = 0; k < rr.n_reac && k < CAM_MAX_REAC; ++k)
        for (int i = 0; i < rx.reac_n[k]; ++i)
            rx.fw[k][rx.reac_at[k][i]] = cfg->mass[rx.reac_at[k][i]] / rx.reac_mass[k];
    for (int i = 0; i < na; ++i) rx.inv_mass[i] = 1.0 / cfg->mass[i];
    rx.inv_nform = rx.n_form > 0 ? 1.0 / rx.n_form : 0.0;
    rx.inv_nbreak = rx.n_break > 0 ? 1.0 / rx.n_break : 0.0;
    rx.inv_terms = rx.n_reac >= 2 ? 2.0 / (rx.n_reac * (rx.n_reac - 1)) : 0.0;
    clear_error();
    std::unique_ptr<cam_engine> e(new cam_engine());
    e->cfg = *cfg;
    e->rx = rx;
    const int L = surf == CAM_SURFACE_CH4H ? 4 : 1;
    e->tpb = kBlock / (L * nb);
    // steps fused per launch: state stays in registers for this many steps
    e->spl = cfg->steps_per_launch > 0 ? cfg->steps_per_launch : 50;
    if (cfg->remove_com) e->spl = 1;
    if (cfg->nm_kernel != CAM_NMK_AUTO) e->nmk = cfg->nm_kernel;
    else e->nmk = fold_available(cfg) ? CAM_NMK_FOLDED : CAM_NMK_VALU;
    // the FP64 CH5 folded kernel is faster at two waves per SIMD, the unfolded
    // step kernels at one (they spill at two)
    const bool ch4h64 = surf == CAM_SURFACE_CH4H && cfg->precision == CAM_PREC_FP64;
    e->wpe = ch4h64 && (cfg->waves_per_eu == 2 || (cfg->waves_per_eu == 0 && e->nmk == CAM_NMK_FOLDED)) ? 2 : 1;
    // windows run fused unless split is asked for or MFMA normal modes are
    // (their fused step has the vector-ALU transform only)
    e->split = cfg->schedule == CAM_SCHED_SPLIT || cfg->ensemble == CAM_ENS_PARENT ||
               (cfg->ensemble == CAM_ENS_UMBRELLA && e->nmk == CAM_NMK_MFMA);
    const bool ok = select_kernels(e.get(), e->wpe, e->nmk, e->kern);
    if (!ok) {
        set_error("nbeads %d not supported", nb);
        return CAM_ERR_UNSUPPORTED;
    }
    CAM_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    CAM_HIP(hipStreamCreateWithFlags(&e->slice_stream, hipStreamNonBlocking));
    CAM_HIP(hipEventCreateWithFlags(&e->slice_fork, hipEventDisableTiming));
    CAM_HIP(hipEventCreateWithFlags(&e->slice_join, hipEventDisableTiming));
    const size_t nstate = (size_t)cfg->ntraj * 3 * na * nb;
    CAM_HIP(hipMalloc(&e->q, nstate * sizeof(double)));
    CAM_HIP(hipMalloc(&e->p, nstate * sizeof(double)));
    CAM_HIP(hipMalloc(&e->g, nstate * sizeof(double)));
    CAM_HIP(hipMalloc(&e->scratch_io, nstate * sizeof(double)));
    CAM_HIP(hipMalloc(&e->vbead, (size_t)cfg->ntraj * nb * sizeof(double)));
    CAM_HIP(hipMalloc(&e->beta, (size_t)cfg->ntraj * sizeof(double)));
    if (cfg->ensemble == CAM_ENS_CHILD) {
        // centroids of one fused launch: the longest schedule the autotuner tries, or the configured one
        const size_t steps = (size_t)std::max(e->spl, 100);
        CAM_HIP(hipMalloc(&e->cbuf, steps * cfg->ntraj * 3 * na * sizeof(double)));
    }
    if (steered) {
        CAM_HIP(hipMalloc(&e->pfail, (size_t)cfg->ntraj * sizeof(int)));
        CAM_HIP(hipMemset(e->pfail, 0, (size_t)cfg->ntraj * sizeof(int)));
    }
    if (cfg->ensemble == CAM_ENS_PARENT) {
        CAM_HIP(hipMalloc(&e->pdxi, (size_t)cfg->ntraj * 3 * na * sizeof(double)));
        CAM_HIP(hipMemset(e->pdxi, 0, (size_t)cfg->ntraj * 3 * na * sizeof(double)));
    }
    if (cfg->ensemble == CAM_ENS_NHC) {
        CAM_HIP(hipMalloc(&e->nhc, (size_t)cfg->ntraj * 4 * sizeof(double)));
        CAM_HIP(hipMemset(e->nhc, 0, (size_t)cfg->ntraj * 4 * sizeof(double)));
    }
    if (cfg->ensemble == CAM_ENS_UMBRELLA) {
        // every window at the reaction's xi_ideal with no force constant until cam_engine_set_windows
        std::vector<double> w(2 * (size_t)cfg->ntraj, 0.0);
        for (int t = 0; t < cfg->ntraj; ++t) w[2 * (size_t)t] = rx.xi_ideal;
        CAM_HIP(hipMalloc(&e->uwin, w.size() * sizeof(double)));
        CAM_HIP(hipMemcpy(e->uwin, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
        CAM_HIP(hipMalloc(&e->ustat, w.size() * sizeof(double)));
        CAM_HIP(hipMemset(e->ustat, 0, w.size() * sizeof(double)));
    }
    CAM_HIP(hipMalloc(&e->weight, (size_t)cfg->ntraj * sizeof(double)));
    CAM_HIP(hipMalloc(&e->den_part, (size_t)cfg->ntraj * sizeof(double)));
    CAM_HIP(hipMalloc(&e->den, sizeof(double)));
    CAM_HIP(hipMalloc(&e->step_base, 2 * sizeof(uint64_t)));
    CAM_HIP(hipMalloc(&e->tfwd, (size_t)nb * nb * sizeof(double)));
    CAM_HIP(hipMalloc(&e->tinv, (size_t)nb * nb * sizeof(double)));
    std::vector<double> fwd, inv;
    nm_matrices(nb, cfg->nm_mode, fwd, inv);
    CAM_HIP(hipMemcpy(e->tfwd, fwd.data(), fwd.size() * sizeof(double), hipMemcpyHostToDevice));
    CAM_HIP(hipMemcpy(e->tinv, inv.data(), inv.size() * sizeof(double), hipMemcpyHostToDevice));
    if (cfg->nm_mode == CAM_NM_REFERENCE && nb >= 4) {
        std::vector<double> fm((size_t)(nb / 2) * (nb / 2));
        fold_matrix(nb, fm.data());
        CAM_HIP(hipMalloc(&e->tfold, fm.size() * sizeof(double)));
        CAM_HIP(hipMemcpy(e->tfold, fm.data(), fm.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    CAM_HIP(hipMemset(e->step_base, 0, 2 * sizeof(uint64_t)));
    CAM_HIP(hipMemset(e->den, 0, sizeof(double)));
    CAM_HIP(hipMemset(e->weight, 0, (size_t)cfg->ntraj * sizeof(double)));
    CAM_HIP(hipEventCreate(&e->ev0));
    CAM_HIP(hipEventCreate(&e->ev1));
    *out = e.release();
    return CAM_OK;
}

int cam_engine_destroy(cam_engine *e)
{
    if (!e) return CAM_OK;
    (void)hipStreamSynchronize(e->stream);
    for (auto &kv : e->graphs) (void)hipGraphExecDestroy(kv.second);
    double *bufs[] = { e->q, e->p, e->g, e->vbead, e->beta, e->tfwd, e->tinv, e->tfold, e->weight, e->den_part,
                       e->den, e->num, e->scratch_io, e->cbuf, e->pdxi, e->uwin, e->ustat, e->nhc };
    for (double *b : bufs)
        if (b) (void)hipFree(b);
    if (e->pfail) (void)hipFree(e->pfail);
    if (e->uhist) (void)hipFree(e->uhist);
    if (e->xi_flag) (void)hipFree(e->xi_flag);
    if (e->step_base) (void)hipFree(e->step_base);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->slice_fork) (void)hipEventDestroy(e->slice_fork);
    if (e->slice_join) (void)hipEventDestroy(e->slice_join);
    if (e->slice_stream) (void)hipStreamDestroy(e->slice_stream);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return CAM_OK;
}

int cam_engine_set_state(cam_engine *e, const double *q, const double *p, const double *beta)
{
    CAM_CHECK_ARG(e && q && p, "null argument");
    CAM_CHECK_ARG(beta || e->cfg.ensemble == CAM_ENS_NVE, "beta is required for this ensemble");
    const int nt = e->cfg.ntraj, nc = 3 * e->cfg.natoms, nb = e->cfg.nbeads;
    const size_t n = (size_t)nt * nc * nb;
    std::vector<double> hb(nt, 1.0);
    if (beta) {
        for (int t = 0; t < nt; ++t) CAM_CHECK_ARG(beta[t] > 0.0, "beta[%d] must be > 0", t);
        std::copy(beta, beta + nt, hb.begin());
    }
    clear_error();
    const int grid = (int)((n + 255) / 256);
    CAM_HIP(hipMemcpyAsync(e->scratch_io, q, n * sizeof(double), hipMemcpyHostToDevice, e->stream));
    hipLaunchKernelGGL(to_device_layout, dim3(grid), dim3(256), 0, e->stream, e->scratch_io, e->q, nt, nc, nb);
    CAM_LAUNCH_CHECK();
    CAM_HIP(hipMemcpyAsync(e->scratch_io, p, n * sizeof(double), hipMemcpyHostToDevice, e->stream));
    hipLaunchKernelGGL(to_device_layout, dim3(grid), dim3(256), 0, e->stream, e->scratch_io, e->p, nt, nc, nb);
    CAM_LAUNCH_CHECK();
    CAM_HIP(hipMemcpyAsync(e->beta, hb.data(), nt * sizeof(double), hipMemcpyHostToDevice, e->stream));
    // a new state starts its Nose-Hoover chains at rest, as mdinit does (src/mdinit.f90:141-145)
    if (e->nhc) CAM_HIP(hipMemsetAsync(e->nhc, 0, (size_t)nt * 4 * sizeof(double), e->stream));
    int rc = run_forces(e);
    if (rc != CAM_OK) return rc;
    CAM_HIP(hipStreamSynchronize(e->stream));
    e->have_state = true;
    e->parent_ready = false;
    return CAM_OK;
}

int cam_engine_get_state(cam_engine *e, double *q, double *p)
{
    CAM_CHECK_ARG(e && (q || p), "null argument");
    CAM_CHECK_ARG(e->have_state, "no state uploaded");
    const int nt = e->cfg.ntraj, nc = 3 * e->cfg.natoms, nb = e->cfg.nbeads;
    const size_t n = (size_t)nt * nc * nb;
    const int grid = (int)((n + 255) / 256);
    clear_error();
    if (q) {
        hipLaunchKernelGGL(from_device_layout, dim3(grid), dim3(256), 0, e->stream, e->q, e->scratch_io, nt, nc, nb);
        CAM_LAUNCH_CHECK();
        CAM_HIP(hipMemcpyAsync(q, e->scratch_io, n * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        CAM_HIP(hipStreamSynchronize(e->stream));
    }
    if (p) {
        hipLaunchKernelGGL(from_device_layout, dim3(grid), dim3(256), 0, e->stream, e->p, e->scratch_io, nt, nc, nb);
        CAM_LAUNCH_CHECK();
        CAM_HIP(hipMemcpyAsync(p, e->scratch_io, n * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        CAM_HIP(hipStreamSynchronize(e->stream));
    }
    return CAM_OK;
}

int cam_engine_get_forces(cam_engine *e, double *dvdq, double *vbead)
{
    CAM_CHECK_ARG(e && (dvdq || vbead), "null argument");
    CAM_CHECK_ARG(e->have_state, "no state uploaded");
    const int nt = e->cfg.ntraj, nc = 3 * e->cfg.natoms, nb = e->cfg.nbeads;
    const size_t n = (size_t)nt * nc * nb;
    clear_error();
    if (dvdq) {
        const int grid = (int)((n + 255) / 256);
        hipLaunchKernelGGL(from_device_layout, dim3(grid), dim3(256), 0, e->stream, e->g, e->scratch_io, nt, nc, nb);
        CAM_LAUNCH_CHECK();
        CAM_HIP(hipMemcpyAsync(dvdq, e->scratch_io, n * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    }
    if (vbead) CAM_HIP(hipMemcpyAsync(vbead, e->vbead, (size_t)nt * nb * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    CAM_HIP(hipStreamSynchronize(e->stream));
    return CAM_OK;
}

int cam_engine_draw_momenta(cam_engine *e, uint64_t draw, int pair)
{
    CAM_CHECK_ARG(e, "null engine");
    CAM_CHECK_ARG(e->have_state, "no state uploaded");
    CAM_CHECK_ARG(!pair || (e->cfg.ntraj % 2) == 0, "antithetic pairs need an even ntraj");
    clear_error();
    const size_t n = (size_t)e->cfg.ntraj * 3 * e->cfg.natoms * e->cfg.nbeads;
    const TrajArgs ta = traj_args(e);
    // draw in the ABI layout order, then convert
    hipLaunchKernelGGL(draw_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, e->stream, e->p, e->beta,
                       e->cfg.ntraj, e->cfg.natoms, e->cfg.nbeads, e->cfg.seed, draw, pair, ta);
    CAM_LAUNCH_CHECK();
    CAM_HIP(hipStreamSynchronize(e->stream));
    return CAM_OK;
}

int cam_engine_constrain(cam_engine *e, int project_p, int *status)
{
    CAM_CHECK_ARG(e, "null engine");
    CAM_CHECK_ARG(e->have_state, "no state uploaded");
    CAM_CHECK_ARG(e->rx.n_form + e->rx.n_break > 0 && e->rx.n_reac >= 2, "no reaction coordinate configured");
    clear_error();
    int *dstat = nullptr;
    CAM_HIP(hipMalloc(&dstat, (size_t)e->cfg.ntraj * sizeof(int)));
    const TrajArgs ta = traj_args(e);
    const int grid = (e->cfg.ntraj + 127) / 128;
    with_natoms(e->cfg.natoms, [&](auto na) {
        hipLaunchKernelGGL(constrain_kernel<decltype(na)::value>, dim3(grid), dim3(128), 0, e->stream, ta, project_p, dstat);
    });
    hipError_t le = hipGetLastError();
    if (le == hipSuccess && status)
        le = hipMemcpyAsync(status, dstat, (size_t)e->cfg.ntraj * sizeof(int), hipMemcpyDeviceToHost, e->stream);
    if (le == hipSuccess) le = hipStreamSynchronize(e->stream);
    (void)hipFree(dstat);
    CAM_HIP(le);
    int rc = run_forces(e);
    if (rc != CAM_OK) return rc;
    CAM_HIP(hipStreamSynchronize(e->stream));
    return CAM_OK;
}

int cam_engine_begin_children(cam_engine *e)
{
    CAM_CHECK_ARG(e, "null engine");
    CAM_CHECK_ARG(e->have_state, "no state uploaded");
    CAM_CHECK_ARG(e->cfg.ensemble == CAM_ENS_CHILD, "begin_children needs ensemble CHILD");
    clear_error();
    const TrajArgs ta = traj_args(e);
    const int grid = (e->cfg.ntraj + 127) / 128;
    with_natoms(e->cfg.natoms, [&](auto na) {
        hipLaunchKernelGGL(child_begin_kernel<decltype(na)::value>, dim3(grid), dim3(128), 0, e->stream, ta, e->weight, e->den_part);
    });
    CAM_LAUNCH_CHECK();
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, e->stream, e->den_part, e->cfg.ntraj, e->den);
    CAM_LAUNCH_CHECK();
    hipLaunchKernelGGL(reset_child_kernel, dim3(1), dim3(1), 0, e->stream, e->step_base);
    CAM_LAUNCH_CHECK();
    CAM_HIP(hipStreamSynchronize(e->stream));
    e->child_steps = 0;
    return CAM_OK;
}

int cam_engine_begin_parent(cam_engine *e, double k_force, double epot_max, int check_after)
{
    CAM_CHECK_ARG(e, "null engine");
    CAM_CHECK_ARG(e->have_state, "no state uploaded");
    CAM_CHECK_ARG(e->cfg.ensemble == CAM_ENS_PARENT, "begin_parent needs ensemble PARENT");
    CAM_CHECK_ARG(std::isfinite(k_force), "k_force must be finite");
    CAM_CHECK_ARG(!std::isnan(epot_max), "epot_max must not be NaN");
    clear_error();
    e->pcheck_max = epot_max;
    e->pcheck_after = check_after;
    drop_graphs(e);                   // the check parameters are arguments of the captured launches
    int rc = run_forces(e);
    if (rc != CAM_OK) return rc;
    const TrajArgs ta = traj_args(e);
    const int pgrid = (e->cfg.ntraj + kParentWaves - 1) / kParentWaves;
    with_natoms(e->cfg.natoms, [&](auto na) {
        hipLaunchKernelGGL(parent_begin_kernel<decltype(na)::value>, dim3(pgrid), dim3(64 * kParentWaves), 0, e->stream, ta, e->g, k_force, e->pdxi, e->pfail);
    });
    CAM_LAUNCH_CHECK();
    hipLaunchKernelGGL(reset_child_kernel, dim3(1), dim3(1), 0, e->stream, e->step_base);   // steps since the start
    CAM_LAUNCH_CHECK();
    CAM_HIP(hipStreamSynchronize(e->stream));
    e->parent_ready = true;
    return CAM_OK;
}

int cam_engine_set_windows(cam_engine *e, const double *xi_ideal, const double *k_force)
{
    CAM_CHECK_ARG(e && xi_ideal && k_force, "null argument");
    CAM_CHECK_ARG(e->cfg.ensemble == CAM_ENS_UMBRELLA, "set_windows needs ensemble UMBRELLA");
    const int nt = e->cfg.ntraj;
    std::vector<double> w(2 * (size_t)nt);
    for (int t = 0; t < nt; ++t) {
        CAM_CHECK_ARG(std::isfinite(xi_ideal[t]) && std::isfinite(k_force[t]), "window %d is not finite", t);
        w[2 * (size_t)t] = xi_ideal[t];
        w[2 * (size_t)t + 1] = k_force[t];
    }
    clear_error();
    CAM_HIP(hipMemcpyAsync(e->uwin, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice, e->stream));
    CAM_HIP(hipStreamSynchronize(e->stream));
    e->parent_ready = false;          // the start's bias depends on the windows
    return CAM_OK;
}

int cam_engine_begin_umbrella(cam_engine *e, int bias, double epot_max, double xi_tol, int check_after)
{
    CAM_CHECK_ARG(e, "null engine");
    CAM_CHECK_ARG(e->have_state, "no state uploaded");
    CAM_CHECK_ARG(e->cfg.ensemble == CAM_ENS_UMBRELLA, "begin_umbrella needs ensemble UMBRELLA");
    CAM_CHECK_ARG(!std::isnan(epot_max) && !std::isnan(xi_tol), "epot_max and xi_tol must not be NaN");
    clear_error();
    e->pcheck_max = epot_max;
    e->pcheck_xi_tol = xi_tol;
    e->pcheck_after = check_after;
    drop_graphs(e);                   // the check parameters are arguments of the captured launches
    int rc = run_forces(e);
    if (rc != CAM_OK) return rc;
    const TrajArgs ta = traj_args(e);
    const int pgrid = (e->cfg.ntraj + kParentWaves - 1) / kParentWaves;
    with_natoms(e->cfg.natoms, [&](auto na) {
        hipLaunchKernelGGL(umbrella_begin_kernel<decltype(na)::value>, dim3(pgrid), dim3(64 * kParentWaves), 0, e->stream, ta, e->g, e->uwin, e->ustat, e->pfail, bias);
    });
    CAM_LAUNCH_CHECK();
    hipLaunchKernelGGL(reset_child_kernel, dim3(1), dim3(1), 0, e->stream, e->step_base);   // steps since the start
    CAM_LAUNCH_CHECK();
    if (e->uhist)
        CAM_HIP(hipMemsetAsync(e->uhist, 0, (size_t)e->cfg.ntraj * e->hbins * sizeof(unsigned), e->stream));
    CAM_HIP(hipStreamSynchronize(e->stream));
    e->ustat_steps = 0;
    e->parent_ready = true;
    return CAM_OK;
}

int cam_engine_set_umbrella_histogram(cam_engine *e, int nbins, double xi_min, double bin_size)
{
    CAM_CHECK_ARG(e, "null engine");
    CAM_CHECK_ARG(e->cfg.ensemble == CAM_ENS_UMBRELLA, "umbrella histograms need ensemble UMBRELLA");
    CAM_CHECK_ARG(nbins >= 0 && nbins <= (1 << 20), "nbins must be in 0..2^20");
    CAM_CHECK_ARG(nbins == 0 || (std::isfinite(xi_min) && std::isfinite(bin_size) && bin_size > 0.0),
                  "xi_min must be finite and bin_size positive");
    clear_error();
    CAM_HIP(hipStreamSynchronize(e->stream));         // no launch may still count into the old histograms
    drop_graphs(e);                                   // the histogram is an argument of the captured launches
    if (e->uhist) CAM_HIP(hipFree(e->uhist));
    e->uhist = nullptr;
    e->hbins = 0;
    if (nbins > 0) {
        const size_t bytes = (size_t)e->cfg.ntraj * nbins * sizeof(unsigned);
        CAM_HIP(hipMalloc(&e->uhist, bytes));
        CAM_HIP(hipMemset(e->uhist, 0, bytes));
        e->hbins = nbins;
        e->hmin = xi_min;
        e->hinv = 1.0 / bin_size;
    }
    return CAM_OK;
}

int cam_engine_umbrella_histogram(cam_engine *e, double *counts)
{
    CAM_CHECK_ARG(e && counts, "null argument");
    CAM_CHECK_ARG(e->uhist, "no umbrella histogram (cam_engine_set_umbrella_histogram)");
    const size_t n = (size_t)e->cfg.ntraj * e->hbins;
    std::vector<unsigned> h(n);
    clear_error();
    CAM_HIP(hipMemcpyAsync(h.data(), e->uhist, n * sizeof(unsigned), hipMemcpyDeviceToHost, e->stream));
    CAM_HIP(hipStreamSynchronize(e->stream));
    for (size_t i = 0; i < n; ++i) counts[i] = (double)h[i];
    return CAM_OK;
}

int cam_engine_umbrella_stats(cam_engine *e, double *average, double *variance)
{
    CAM_CHECK_ARG(e && average && variance, "null argument");
    CAM_CHECK_ARG(e->cfg.ensemble == CAM_ENS_UMBRELLA, "umbrella_stats needs ensemble UMBRELLA");
    CAM_CHECK_ARG(e->parent_ready && e->ustat_steps > 0, "no window steps since cam_engine_begin_umbrella");
    const int nt = e->cfg.ntraj;
    std::vector<double> s(2 * (size_t)nt);
    clear_error();
    CAM_HIP(hipMemcpyAsync(s.data(), e->ustat, s.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    CAM_HIP(hipStreamSynchronize(e->stream));
    const double n = (double)e->ustat_steps;
    for (int t = 0; t < nt; ++t) {
        // src/calc_rate.f90:1655-1657
        average[t] = s[2 * (size_t)t] / n;
        variance[t] = s[2 * (size_t)t + 1] / n - average[t] * average[t];
    }
    return CAM_OK;
}

static int get_graph(cam_engine *e, int nsteps, hipGraphExec_t *out)
{
    if (e->cfg.ensemble == CAM_ENS_CHILD) {
        int rc = ensure_child_buffers(e, nsteps, e->child_steps + nsteps);
        if (rc != CAM_OK) return rc;
    }
    auto it = e->graphs.find(nsteps);
    if (it == e->graphs.end()) {
        CAM_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_run(e, nsteps, true);
        hipGraph_t graph = nullptr;
        hipError_t ce = hipStreamEndCapture(e->stream, &graph);
        if (rc != CAM_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        CAM_HIP(ce);
        hipGraphExec_t exec;
        hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        CAM_HIP(ie);
        it = e->graphs.emplace(nsteps, exec).first;
    }
    *out = it->second;
    return CAM_OK;
}

int cam_engine_prepare(cam_engine *e, int nsteps)
{
    CAM_CHECK_ARG(e, "null engine");
    CAM_CHECK_ARG(nsteps >= 1, "nsteps must be >= 1");
    CAM_CHECK_ARG(e->have_state, "no state uploaded");
    clear_error();
    hipGraphExec_t g;
    return get_graph(e, nsteps, &g);
}

int cam_engine_run(cam_engine *e, int nsteps)
{
    CAM_CHECK_ARG(e, "null engine");
    CAM_CHECK_ARG(nsteps >= 0, "nsteps must be >= 0");
    CAM_CHECK_ARG(e->have_state, "no state uploaded");
    CAM_CHECK_ARG(e->cfg.ensemble != CAM_ENS_PARENT || e->parent_ready,
                  "a PARENT engine needs cam_engine_begin_parent after cam_engine_set_state");
    CAM_CHECK_ARG(e->cfg.ensemble != CAM_ENS_UMBRELLA || e->parent_ready,
                  "an UMBRELLA engine needs cam_engine_begin_umbrella after cam_engine_set_state");
    clear_error();
    if (nsteps == 0) return CAM_OK;
    // Long runs replay a graph of at most kGraphLaunches step launches (and
    // one graph of the remainder) instead of capturing every step into one
    // graph: a split-schedule step is three or four launches, a fused launch
    // spl steps (one, plus transrot, when the centre of mass is removed).
    const int steps_per_graph = (e->split ? 1 : e->spl) * kGraphLaunches;
    const int chunk = std::min(nsteps, steps_per_graph);
    const int reps = nsteps / chunk, tail = nsteps % chunk;
    if (e->cfg.ensemble == CAM_ENS_CHILD) {
        int rc = ensure_child_buffers(e, chunk, e->child_steps + nsteps);
        if (rc != CAM_OK) return rc;
    }
    hipGraphExec_t g = nullptr, gt = nullptr;
    int rc = get_graph(e, chunk, &g);
    if (rc == CAM_OK && tail) rc = get_graph(e, tail, &gt);
    if (rc != CAM_OK) return rc;
    CAM_HIP(hipEventRecord(e->ev0, e->stream));
    for (int r = 0; r < reps; ++r) CAM_HIP(hipGraphLaunch(g, e->stream));
    if (tail) CAM_HIP(hipGraphLaunch(gt, e->stream));
    CAM_HIP(hipEventRecord(e->ev1, e->stream));
    e->child_steps += nsteps;
    e->ustat_steps += nsteps;
    // the step launches just enqueued: per graph, ceil(steps / spl) fused
    // launches (the last one of a graph may be shorter) or one per split step
    const int per = e->split ? 1 : e->spl;
    e->last_launches = reps * ((chunk + per - 1) / per) + (tail ? (tail + per - 1) / per : 0);
    e->last_steps = nsteps;
    e->last_ms = -1.0;
    return CAM_OK;
}

int cam_engine_sync(cam_engine *e)
{
    CAM_CHECK_ARG(e, "null engine");
    clear_error();
    CAM_HIP(hipStreamSynchronize(e->stream));
    if (e->last_ms < 0.0) {
        float ms = 0.f;
        CAM_HIP(hipEventElapsedTime(&ms, e->ev0, e->ev1));
        e->last_ms = ms;
    }
    return CAM_OK;
}

int cam_engine_kappa(cam_engine *e, double *num, double *den, int nsteps)
{
    CAM_CHECK_ARG(e && num && den, "null argument");
    CAM_CHECK_ARG(e->cfg.ensemble == CAM_ENS_CHILD, "kappa needs ensemble CHILD");
    CAM_CHECK_ARG(nsteps >= 0 && nsteps <= e->num_cap, "nsteps exceeds the child steps recorded (%d)", e->num_cap);
    clear_error();
    if (nsteps) CAM_HIP(hipMemcpyAsync(num, e->num, (size_t)nsteps * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    CAM_HIP(hipMemcpyAsync(den, e->den, sizeof(double), hipMemcpyDeviceToHost, e->stream));
    CAM_HIP(hipStreamSynchronize(e->stream));
    return CAM_OK;
}

int cam_engine_observables(cam_engine *e, double *epot, double *ekin, double *espring, double *xi, int *flags)
{
    CAM_CHECK_ARG(e && epot && ekin && espring && xi && flags, "null argument");
    CAM_CHECK_ARG(e->have_state, "no state uploaded");
    clear_error();
    const int nt = e->cfg.ntraj;
    double *buf = nullptr;
    int *fl = nullptr;
    CAM_HIP(hipMalloc(&buf, (size_t)4 * nt * sizeof(double)));
    hipError_t err = hipMalloc(&fl, (size_t)nt * sizeof(int));
    if (err == hipSuccess) {
        const TrajArgs ta = traj_args(e);
        const int with_xi = e->cfg.ensemble != CAM_ENS_NVE && e->cfg.ensemble != CAM_ENS_ANDERSEN;
        hipLaunchKernelGGL(observe_kernel, dim3((nt + 127) / 128), dim3(128), 0, e->stream, ta, e->pfail,
                           buf, buf + nt, buf + 2 * nt, buf + 3 * nt, fl, with_xi);
        err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipMemcpyAsync(epot, buf, nt * sizeof(double), hipMemcpyDeviceToHost, e->stream);
    if (err == hipSuccess) err = hipMemcpyAsync(ekin, buf + nt, nt * sizeof(double), hipMemcpyDeviceToHost, e->stream);
    if (err == hipSuccess) err = hipMemcpyAsync(espring, buf + 2 * nt, nt * sizeof(double), hipMemcpyDeviceToHost, e->stream);
    if (err == hipSuccess) err = hipMemcpyAsync(xi, buf + 3 * nt, nt * sizeof(double), hipMemcpyDeviceToHost, e->stream);
    if (err == hipSuccess) err = hipMemcpyAsync(flags, fl, nt * sizeof(int), hipMemcpyDeviceToHost, e->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
    (void)hipFree(buf);
    if (fl) (void)hipFree(fl);
    CAM_HIP(err);
    return CAM_OK;
}

int cam_engine_device_state(cam_engine *e, double **q, double **p, double **dvdq, void **stream)
{
    CAM_CHECK_ARG(e, "null engine");
    if (q) *q = e->q;
    if (p) *p = e->p;
    if (dvdq) *dvdq = e->g;
    if (stream) *stream = (void *)e->stream;
    return CAM_OK;
}

// slices: the trials launch as the engine's own runs do (true) or as a
// ragged batch's bucket does (false: no trajectory slices).
static int autotune(cam_engine *e, int trial_steps, cam_tuning *out, bool slices);

int cam_engine_autotune(cam_engine *e, int trial_steps, cam_tuning *out)
{
    return autotune(e, trial_steps, out, true);
}

static int autotune(cam_engine *e, int trial_steps, cam_tuning *out, bool slices)
{
    CAM_CHECK_ARG(e, "null engine");
    CAM_CHECK_ARG(e->have_state, "no state uploaded");
    CAM_CHECK_ARG(trial_steps >= 1 && trial_steps <= 100000, "trial_steps must be in 1..100000");
    CAM_CHECK_ARG(e->cfg.ensemble != CAM_ENS_PARENT || e->parent_ready,
                  "a PARENT engine needs cam_engine_begin_parent before autotune");
    CAM_CHECK_ARG(e->cfg.ensemble != CAM_ENS_UMBRELLA || e->parent_ready,
                  "an UMBRELLA engine needs cam_engine_begin_umbrella before autotune");
    clear_error();
    const cam_config &c = e->cfg;
    const int nb = c.nbeads;
    std::vector<int> nmks, spls, wpes;
    if (c.waves_per_eu != 0 || c.surface != CAM_SURFACE_CH4H || c.precision != CAM_PREC_FP64) wpes.push_back(e->wpe);
    else wpes = { 1, 2 };
    if (c.nm_kernel != CAM_NMK_AUTO) nmks.push_back(c.nm_kernel);
    else {
        nmks.push_back(CAM_NMK_VALU);
        if (mfma_bead_count(nb)) nmks.push_back(CAM_NMK_MFMA);
        if (fold_available(&c)) nmks.push_back(CAM_NMK_FOLDED);
    }
    if (c.remove_com) spls.push_back(1);
    else if (c.steps_per_launch > 0) spls.push_back(c.steps_per_launch);
    else
        for (int v : { 10, 25, 50, 100 }) spls.push_back(v);
    // Trials run on a copy of the state with their own flag rows and step
    // counters, so the engine's trajectories, kappa sums and draws are untouched.
    const size_t nstate = (size_t)c.ntraj * 3 * c.natoms * nb;
    double *buf = nullptr;
    uint8_t *flags = nullptr;
    uint64_t *ctr = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    int rc = CAM_OK;
    struct Trial { int nmk, wpe, spl, split; double us; };
    std::vector<Trial> cand, trials;
    const bool parent = c.ensemble == CAM_ENS_PARENT, window = c.ensemble == CAM_ENS_UMBRELLA;
    const bool nhc = c.ensemble == CAM_ENS_NHC;
    const bool vfused = window || nhc;          // a fused step with the VALU transform and one budget only
    const bool try_fused = c.schedule != CAM_SCHED_SPLIT && !parent;
    const bool try_split = c.schedule != CAM_SCHED_FUSED && !c.remove_com && !nhc;
    for (int nmk : nmks) {
        if (nmk == CAM_NMK_FOLDED) {       // fused only
            for (int wpe : wpes)
                for (int spl : spls) cand.push_back({ nmk, wpe, spl, 0, 0.0 });
            continue;
        }
        if (try_fused && !(vfused && nmk != CAM_NMK_VALU))
            for (int wpe : wpes)
                if (!vfused || wpe == 1)
                    for (int spl : spls) cand.push_back({ nmk, wpe, spl, 0, 0.0 });
        if (try_split) cand.push_back({ nmk, 1, 1, 1, 0.0 });
    }
    auto fail = [&](hipError_t err) {
        set_error("HIP error in autotune: %s", hipGetErrorString(err));
        rc = CAM_ERR_DEVICE;
    };
    // PARENT: a copy of the constraint directions; UMBRELLA: scratch sums;
    // NHC: scratch chains; all: scratch flags
    const size_t npd = parent ? (size_t)c.ntraj * 3 * c.natoms
                              : window ? 2 * (size_t)c.ntraj : nhc ? 4 * (size_t)c.ntraj : 0;
    hipError_t err = hipMalloc(&buf, (3 * nstate + (size_t)c.ntraj * nb + npd + (size_t)c.ntraj) * sizeof(double));
    if (err == hipSuccess) err = hipMalloc(&flags, (size_t)trial_steps * c.ntraj);
    if (err == hipSuccess) err = hipMalloc(&ctr, 2 * sizeof(uint64_t));
    if (err == hipSuccess) err = hipEventCreate(&t0);
    if (err == hipSuccess) err = hipEventCreate(&t1);
    if (err != hipSuccess) fail(err);
    for (const Trial &tr : cand) {
        if (rc != CAM_OK) break;
        Kernels kn;
        if (!select_kernels(e, tr.wpe, tr.nmk, kn)) continue;
        StepArgs a = step_args(e);
        a.q = buf;
        a.p = buf + nstate;
        a.g = buf + 2 * nstate;
        a.vbead = buf + 3 * nstate;
        a.xi_flag = flags;
        a.step_base = ctr;
        if (parent) {
            a.pdxi = buf + 3 * nstate + (size_t)c.ntraj * nb;
            a.pfail = reinterpret_cast<int *>(a.pdxi + npd);
        }
        if (window) {
            a.ustat = buf + 3 * nstate + (size_t)c.ntraj * nb;
            a.uhist = nullptr;                      // trials count nothing
            a.pfail = reinterpret_cast<int *>(a.ustat + npd);
        }
        if (nhc) a.nhc = buf + 3 * nstate + (size_t)c.ntraj * nb;
        err = hipMemsetAsync(ctr, 0, 2 * sizeof(uint64_t), e->stream);
        if (err == hipSuccess && parent)
            err = hipMemcpyAsync(a.pdxi, e->pdxi, npd * sizeof(double), hipMemcpyDeviceToDevice, e->stream);
        if (err == hipSuccess && nhc)
            err = hipMemcpyAsync(a.nhc, e->nhc, npd * sizeof(double), hipMemcpyDeviceToDevice, e->stream);
        if (err == hipSuccess) err = hipMemcpyAsync(a.q, e->q, nstate * sizeof(double), hipMemcpyDeviceToDevice, e->stream);
        if (err == hipSuccess) err = hipMemcpyAsync(a.p, e->p, nstate * sizeof(double), hipMemcpyDeviceToDevice, e->stream);
        if (err == hipSuccess) err = hipMemcpyAsync(a.g, e->g, nstate * sizeof(double), hipMemcpyDeviceToDevice, e->stream);
        // one untimed launch sequence, then the timed trial
        if (err == hipSuccess)
            err = launch_steps(e, kn, tr.split != 0, tr.spl, a, std::min(tr.spl, trial_steps), nullptr, slices);
        if (err == hipSuccess) err = hipEventRecord(t0, e->stream);
        if (err == hipSuccess) err = launch_steps(e, kn, tr.split != 0, tr.spl, a, trial_steps, nullptr, slices);
        if (err == hipSuccess) err = hipEventRecord(t1, e->stream);
        if (err == hipSuccess) err = hipEventSynchronize(t1);
        float ms = 0.f;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, t0, t1);
        if (err != hipSuccess) {
            fail(err);
            break;
        }
        Trial done = tr;
        done.us = 1e3 * ms / trial_steps;
        trials.push_back(done);
    }
    if (buf) (void)hipFree(buf);
    if (flags) (void)hipFree(flags);
    if (ctr) (void)hipFree(ctr);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (rc != CAM_OK) return rc;
    CAM_CHECK_ARG(!trials.empty(), "no candidate schedule");
    size_t best = 0;
    for (size_t i = 1; i < trials.size(); ++i)
        if (trials[i].us < trials[best].us) best = i;
    Kernels kn;
    e->wpe = trials[best].wpe;
    select_kernels(e, e->wpe, trials[best].nmk, kn);
    e->nmk = trials[best].nmk;
    e->kern = kn;
    e->spl = trials[best].spl;
    e->split = trials[best].split != 0;
    drop_graphs(e);
    if (out) {
        std::memset(out, 0, sizeof *out);
        out->n = (int)std::min<size_t>(trials.size(), CAM_MAX_TRIALS);
        out->chosen = (int)best;
        for (int i = 0; i < out->n; ++i) {
            out->nm_kernel[i] = trials[i].nmk;
            out->steps_per_launch[i] = trials[i].spl;
            out->waves_per_eu[i] = trials[i].wpe;
            out->split[i] = trials[i].split;
            out->us_per_step[i] = trials[i].us;
        }
    }
    return CAM_OK;
}

// ---------------------------------------------------------------------------
// Ragged batches: one engine per (surface, nbeads) bucket.  A run of the
// batch is one captured graph: forked from the batch stream, every bucket's
// steps on its own stream (branches in decreasing work estimate), joined
// back; the graph is replayed like an engine's and captured again when an
// engine dropped its own graphs (buffers or launch arguments changed).
// ---------------------------------------------------------------------------
struct cam_batch {
    std::vector<cam_engine *> eng;
    std::vector<int> order;            // launch order, decreasing work estimate
    std::vector<hipEvent_t> done;
    hipStream_t stream = nullptr;      // fork and join stream
    hipEvent_t start = nullptr, stop = nullptr, fork = nullptr;
    std::map<int, hipGraphExec_t> graphs;   // by steps per replay
    std::vector<unsigned long> gens;        // the engines' graph generations the graphs were captured at
    bool pending = false;
    double last_ms = 0.0;
};

static void batch_drop_graphs(cam_batch *b)
{
    if (b->graphs.empty()) return;
    (void)hipStreamSynchronize(b->stream);
    for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
}

int cam_batch_destroy(cam_batch *b)
{
    if (!b) return CAM_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (cam_engine *e : b->eng) cam_engine_destroy(e);
    for (hipEvent_t ev : b->done)
        if (ev) (void)hipEventDestroy(ev);
    if (b->start) (void)hipEventDestroy(b->start);
    if (b->stop) (void)hipEventDestroy(b->stop);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return CAM_OK;
}

int cam_batch_create(const cam_config *cfgs, int nbuckets, cam_batch **out)
{
    CAM_CHECK_ARG(cfgs && out, "null argument");
    CAM_CHECK_ARG(nbuckets >= 1 && nbuckets <= 1024, "nbuckets must be in 1..1024 (got %d)", nbuckets);
    *out = nullptr;
    cam_batch *b = new cam_batch();
    std::vector<double> work(nbuckets);
    for (int i = 0; i < nbuckets; ++i) {
        cam_engine *e = nullptr;
        const int rc = cam_engine_create(&cfgs[i], &e);
        if (rc != CAM_OK) {
            cam_batch_destroy(b);
            return rc;
        }
        b->eng.push_back(e);
        // The buckets run at once on one chip: a folded one-lane-per-bead
        // kernel with two column buffers (one block per CU) leaves the
        // registers of each CU's SIMDs to the other buckets' blocks (the
        // ragged H3 + CH5 batch: 4.1e9 against 3.8e9 bead-force-evals/s
        // with one buffer).
        e->fold_nbuf = 2;
        select_kernels(e, e->wpe, e->nmk, e->kern);
        // flops of one step with elementary functions weighted 20 flops each
        const double bead = cfgs[i].surface == CAM_SURFACE_H3     ? 1438.0 + 20.0 * 22
                            : cfgs[i].surface == CAM_SURFACE_OH3  ? 420.0 + 20.0 * 26
                            : cfgs[i].surface == CAM_SURFACE_BRH2 ? 2200.0 + 20.0 * 100
                            : cfgs[i].surface == CAM_SURFACE_O3   ? 2000.0 + 20.0 * 40
                                                                 : 12707.0 + 20.0 * 321;
        work[i] = (double)cfgs[i].ntraj * cfgs[i].nbeads * bead;
    }
    b->order.resize(nbuckets);
    for (int i = 0; i < nbuckets; ++i) b->order[i] = i;
    std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) { return work[x] > work[y]; });
    b->done.assign(nbuckets, nullptr);
    hipError_t err = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipEventCreate(&b->start);
    if (err == hipSuccess) err = hipEventCreate(&b->stop);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    for (int i = 0; err == hipSuccess && i < nbuckets; ++i) err = hipEventCreateWithFlags(&b->done[i], hipEventDisableTiming);
    if (err != hipSuccess) {
        set_error("HIP error: %s", hipGetErrorString(err));
        cam_batch_destroy(b);
        return CAM_ERR_DEVICE;
    }
    *out = b;
    return CAM_OK;
}

int cam_batch_engine(cam_batch *b, int i, cam_engine **eng)
{
    CAM_CHECK_ARG(b && eng, "null argument");
    CAM_CHECK_ARG(i >= 0 && i < (int)b->eng.size(), "bucket %d out of range", i);
    *eng = b->eng[i];
    return CAM_OK;
}

int cam_batch_autotune(cam_batch *b, int trial_steps)
{
    CAM_CHECK_ARG(b, "null batch");
    for (cam_engine *e : b->eng) {
        const int rc = autotune(e, trial_steps, nullptr, false);
        if (rc != CAM_OK) return rc;
    }
    return CAM_OK;
}

// The graph of one replay of nsteps steps of every bucket.
static int batch_graph(cam_batch *b, int nsteps, hipGraphExec_t *out)
{
    bool stale = b->gens.size() != b->eng.size();
    for (size_t i = 0; !stale && i < b->eng.size(); ++i) stale = b->gens[i] != b->eng[i]->graph_gen;
    if (stale) {
        batch_drop_graphs(b);
        b->gens.resize(b->eng.size());
        for (size_t i = 0; i < b->eng.size(); ++i) b->gens[i] = b->eng[i]->graph_gen;
    }
    auto it = b->graphs.find(nsteps);
    if (it == b->graphs.end()) {
        CAM_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = CAM_OK;
        hipError_t err = hipEventRecord(b->fork, b->stream);
        for (int i : b->order) {
            if (err != hipSuccess || rc != CAM_OK) break;
            cam_engine *e = b->eng[i];
            err = hipStreamWaitEvent(e->stream, b->fork, 0);
            if (err != hipSuccess) break;
            rc = enqueue_run(e, nsteps, false);       // the buckets fill the chip together
            // the branch is joined even after a failure, so the capture ends cleanly
            hipError_t je = hipEventRecord(b->done[i], e->stream);
            if (je == hipSuccess) je = hipStreamWaitEvent(b->stream, b->done[i], 0);
            if (err == hipSuccess) err = je;
        }
        hipGraph_t graph = nullptr;
        const hipError_t ce = hipStreamEndCapture(b->stream, &graph);
        if (rc != CAM_OK || err != hipSuccess || ce != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            if (rc != CAM_OK) return rc;
            CAM_HIP(err != hipSuccess ? err : ce);
        }
        hipGraphExec_t exec;
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        CAM_HIP(ie);
        it = b->graphs.emplace(nsteps, exec).first;
    }
    *out = it->second;
    return CAM_OK;
}

int cam_batch_run(cam_batch *b, int nsteps)
{
    CAM_CHECK_ARG(b, "null batch");
    CAM_CHECK_ARG(nsteps >= 0, "nsteps must be >= 0");
    for (cam_engine *e : b->eng) {
        CAM_CHECK_ARG(e->have_state, "no state uploaded");
        CAM_CHECK_ARG(e->cfg.ensemble != CAM_ENS_PARENT && e->cfg.ensemble != CAM_ENS_UMBRELLA,
                      "batches run NVE, Andersen, Nose-Hoover and child engines");
    }
    clear_error();
    if (nsteps == 0) return CAM_OK;
    // replays of at most kGraphLaunches step launches of every bucket
    int chunk = nsteps;
    for (cam_engine *e : b->eng) chunk = std::min(chunk, (e->split ? 1 : e->spl) * kGraphLaunches);
    const int reps = nsteps / chunk, tail = nsteps % chunk;
    for (cam_engine *e : b->eng)
        if (e->cfg.ensemble == CAM_ENS_CHILD && (chunk > e->flag_cap || e->child_steps + nsteps > e->num_cap)) {
            // the buffers grow: nothing may still use them
            CAM_HIP(hipStreamSynchronize(b->stream));
            const int rc = ensure_child_buffers(e, chunk, e->child_steps + nsteps);
            if (rc != CAM_OK) return rc;
        }
    hipGraphExec_t g = nullptr, gt = nullptr;
    int rc = batch_graph(b, chunk, &g);
    if (rc == CAM_OK && tail) rc = batch_graph(b, tail, &gt);
    if (rc != CAM_OK) return rc;
    CAM_HIP(hipEventRecord(b->start, b->stream));
    for (cam_engine *e : b->eng) CAM_HIP(hipEventRecord(e->ev0, b->stream));
    for (int r = 0; r < reps; ++r) CAM_HIP(hipGraphLaunch(g, b->stream));
    if (tail) CAM_HIP(hipGraphLaunch(gt, b->stream));
    CAM_HIP(hipEventRecord(b->stop, b->stream));
    for (cam_engine *e : b->eng) {
        CAM_HIP(hipEventRecord(e->ev1, b->stream));
        e->child_steps += nsteps;
        e->ustat_steps += nsteps;
        const int per = e->split ? 1 : e->spl;
        e->last_launches = reps * ((chunk + per - 1) / per) + (tail ? (tail + per - 1) / per : 0);
        e->last_steps = nsteps;
        e->last_ms = -1.0;
    }
    b->pending = true;
    return CAM_OK;
}

int cam_batch_sync(cam_batch *b, double *ms)
{
    CAM_CHECK_ARG(b, "null batch");
    clear_error();
    CAM_HIP(hipStreamSynchronize(b->stream));
    if (b->pending) {
        float t = 0.f;
        CAM_HIP(hipEventElapsedTime(&t, b->start, b->stop));
        b->last_ms = t;
        b->pending = false;
        for (cam_engine *e : b->eng) {
            const int rc = cam_engine_sync(e);
            if (rc != CAM_OK) return rc;
        }
    }
    if (ms) *ms = b->last_ms;
    return CAM_OK;
}

int cam_engine_last_timing(cam_engine *e, double *ms_per_launch, int *steps_per_launch, int *launches)
{
    CAM_CHECK_ARG(e, "null engine");
    CAM_CHECK_ARG(e->last_ms >= 0.0 && e->last_launches > 0, "no completed run (call cam_engine_sync first)");
    if (ms_per_launch) *ms_per_launch = e->last_ms / e->last_launches;
    if (steps_per_launch) *steps_per_launch = (int)std::min<long>(e->split ? 1 : e->spl, e->last_steps);
    if (launches) *launches = e->last_launches;
    return CAM_OK;
}

}  // extern "C"
