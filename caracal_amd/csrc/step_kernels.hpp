// Step kernels of the trajectory-batch engine (the fused step with its
// normal-mode transforms, the split schedule's normal-mode and force kernels,
// the initial-force kernel) and the helpers they share with the engine's host
// code.  Each translation unit that includes this header instantiates the
// kernels of the surfaces it picks (engine_h3.hip, engine_ch4h.hip,
// engine_fp64.hip), so that they compile in parallel; engine.hip uses the
// helpers and kernels that take no surface.
#pragma once
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "philox.hpp"
#include "step_common.hpp"

namespace cam {
namespace {

// Recrossing parents and umbrella windows: trajectories per block of their
// per-trajectory kernels, step launches per captured graph (a split-schedule
// step is one launch sequence; a fused launch advances steps_per_launch
// steps), and the bits of the per-trajectory flags (a constraint did not
// converge since the start; the trajectory check failed; this step's
// constraint did not converge).
constexpr int kParentWaves = 4;
constexpr int kGraphLaunches = 1000;
constexpr int kPfConstraint = 1, kPfCheck = 2, kPfThisStep = 4;
// Folded launches of at least 2 x kSliceMinBlocks blocks (the CU count) are
// issued as two trajectory slices on two streams (launch_steps).
constexpr int kSliceMinBlocks = 256;

// xi of the centroid without its gradient, for the per-step child flags
// (calc_xi BIMOLEC mode 2, src/calc_xi.f90:63-213).  One lane evaluates it
// while its block waits, so everything it reads is in LDS or registers:
// wcom[k][at] = mass[at] / M_k for the atoms of fragment k (0 otherwise), and
// the bond loops are unrolled to their maximum under run-time guards.  The
// fragment centres of mass are sums of weighted positions, which differs from
// calc_xi's sum of mass * position / M_k only in rounding.
template <int NA>
__device__ double calc_xi_value(const Rx &rx, const double (*wcom)[NA], const double *c)
{
    double s1 = 0.0;
#pragma unroll
    for (int i = 0; i < CAM_MAX_BONDS; ++i)
        if (i < rx.n_break) {
            const int a = rx.brk[i][0], b = rx.brk[i][1];
            const double dx = c[3 * a] - c[3 * b], dy = c[3 * a + 1] - c[3 * b + 1], dz = c[3 * a + 2] - c[3 * b + 2];
            s1 = s1 + (sqrt(dx * dx + dy * dy + dz * dz) - rx.break_ref[i]) / (double)rx.n_break;
        }
#pragma unroll
    for (int i = 0; i < CAM_MAX_BONDS; ++i)
        if (i < rx.n_form) {
            const int a = rx.form[i][0], b = rx.form[i][1];
            const double dx = c[3 * a] - c[3 * b], dy = c[3 * a + 1] - c[3 * b + 1], dz = c[3 * a + 2] - c[3 * b + 2];
            s1 = s1 - (sqrt(dx * dx + dy * dy + dz * dz) - rx.form_ref[i]) / (double)rx.n_form;
        }
    double com[CAM_MAX_REAC][3];
#pragma unroll
    for (int k = 0; k < CAM_MAX_REAC; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int at = 0; at < NA; ++at) acc = fma(wcom[k][at], c[3 * at + j], acc);
            com[k][j] = acc;
        }
    double s0 = 0.0;
    const int terms = (rx.n_reac * rx.n_reac - rx.n_reac) / 2;
#pragma unroll
    for (int i = 0; i < CAM_MAX_REAC; ++i)
#pragma unroll
        for (int j = i + 1; j < CAM_MAX_REAC; ++j)
            if (j < rx.n_reac) {
                const double dx = com[j][0] - com[i][0], dy = com[j][1] - com[i][1], dz = com[j][2] - com[i][2];
                s0 = s0 + (rx.r_inf - sqrt(dx * dx + dy * dy + dz * dz));
            }
    s0 = s0 / (double)terms;
    return rx.xi_ideal * s1 + (1.0 - rx.xi_ideal) * s0;
}

// The helpers below run on values in registers: every index into a private
// array is a compile-time constant (loops over atoms, bonds and fragments are
// unrolled to their maxima under run-time guards), so nothing goes to scratch
// memory.  An atom named by the reaction is read or updated by blending over
// all atoms with a 0/1 weight: as arithmetic, the compiler cannot fold it back
// into an indexed access of the array (a conditional copy it can, and then
// the array lives in scratch).  The blends are exact for finite values.

// v[3*a .. 3*a+2] for a run-time atom index a.
template <int NA>
__device__ inline void gather3(const double *v, int a, double *out)
{
    out[0] = out[1] = out[2] = 0.0;
#pragma unroll
    for (int at = 0; at < NA; ++at) {
        const double f = at == a ? 1.0 : 0.0;
        out[0] = fma(f, v[3 * at], out[0]);
        out[1] = fma(f, v[3 * at + 1], out[1]);
        out[2] = fma(f, v[3 * at + 2], out[2]);
    }
}

// v[3*a + d] += w[d] for a run-time atom index a.
template <int NA>
__device__ inline void scatter_add3(double *v, int a, const double *w)
{
#pragma unroll
    for (int at = 0; at < NA; ++at) {
        const double f = at == a ? 1.0 : 0.0;
        v[3 * at] = fma(f, w[0], v[3 * at]);
        v[3 * at + 1] = fma(f, w[1], v[3 * at + 1]);
        v[3 * at + 2] = fma(f, w[2], v[3 * at + 2]);
    }
}

// x[lane] for lane < N (0 beyond), by selects over the values.
template <int N>
__device__ inline double pick_lane(const double *x, int lane)
{
    double v = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) v = lane == i ? x[i] : v;
    return v;
}

// The two dividing surfaces of calc_xi BIMOLEC (src/calc_xi.f90:63-282) at
// c (3*natoms, atom-major): s1 from the forming and breaking bonds, s0 from
// the distances between the reactant centres of mass, and with GRAD their
// gradients.  Both modes of xi combine these.  Divisions by bond counts,
// term counts and reactant masses are multiplications by the reciprocals and
// fragment weights the reaction carries, and the centres of mass sum over
// all atoms with those weights (rounding differences from the reference's
// sums only).
template <int NA, bool GRAD>
__device__ void xi_surfaces(const Rx &rx, const double *c, double &s0, double &s1, double *ds0, double *ds1)
{
    if constexpr (GRAD) {
#pragma unroll
        for (int i = 0; i < 3 * NA; ++i) ds1[i] = ds0[i] = 0.0;
    }
    s1 = 0.0;
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {                  // breaking bonds (+), then forming bonds (-)
#pragma unroll
        for (int i = 0; i < CAM_MAX_BONDS; ++i) {
            const int nbond = pass == 0 ? rx.n_break : rx.n_form;
            if (i >= nbond) continue;
            const int a = pass == 0 ? rx.brk[i][0] : rx.form[i][0], b = pass == 0 ? rx.brk[i][1] : rx.form[i][1];
            const double inv_n = pass == 0 ? rx.inv_nbreak : rx.inv_nform;
            double ca[3], cb[3], R[3];
            gather3<NA>(c, a, ca);
            gather3<NA>(c, b, cb);
#pragma unroll
            for (int d = 0; d < 3; ++d) R[d] = ca[d] - cb[d];
            const double r = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
            if (pass == 0) s1 = s1 + (r - rx.break_ref[i]) * inv_n;
            else s1 = s1 - (r - rx.form_ref[i]) * inv_n;
            if constexpr (GRAD) {
                const double f = (pass == 0 ? 1.0 : -1.0) * inv_n / r;
                double w[3], mw[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    w[d] = R[d] * f;
                    mw[d] = -w[d];
                }
                scatter_add3<NA>(ds1, a, w);
                scatter_add3<NA>(ds1, b, mw);
            }
        }
    }
    double com[CAM_MAX_REAC][3];
#pragma unroll
    for (int k = 0; k < CAM_MAX_REAC; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int at = 0; at < NA; ++at) acc = fma(rx.fw[k][at], c[3 * at + j], acc);
            com[k][j] = acc;
        }
    s0 = 0.0;
#pragma unroll
    for (int i = 0; i < CAM_MAX_REAC; ++i)
#pragma unroll
        for (int j = i + 1; j < CAM_MAX_REAC; ++j) {
            if (j >= rx.n_reac) continue;
            double red[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) red[d] = com[j][d] - com[i][d];
            const double r = sqrt(red[0] * red[0] + red[1] * red[1] + red[2] * red[2]);
            s0 = s0 + (rx.r_inf - r);
            if constexpr (GRAD) {
                const double f = rx.inv_terms / r;
#pragma unroll
                for (int at = 0; at < NA; ++at) {
                    const double wij = rx.fw[i][at] - rx.fw[j][at];       // an atom is in one reactant at most
#pragma unroll
                    for (int d = 0; d < 3; ++d) ds0[3 * at + d] = fma(red[d] * f, wij, ds0[3 * at + d]);
                }
            }
        }
    s0 = s0 * rx.inv_terms;
}

// xi of the centroid (calc_xi BIMOLEC mode 2, src/calc_xi.f90:209-211,
// :280-281), c: centroid (3*natoms, atom-major).  Optionally the gradient
// dxi (3*natoms).
template <int NA>
__device__ double calc_xi(const Rx &rx, const double *c, double *dxi)
{
    double s0, s1;
    if (!dxi) {
        xi_surfaces<NA, false>(rx, c, s0, s1, nullptr, nullptr);
        return rx.xi_ideal * s1 + (1.0 - rx.xi_ideal) * s0;
    }
    double ds0[3 * NA], ds1[3 * NA];
    xi_surfaces<NA, true>(rx, c, s0, s1, ds0, ds1);
#pragma unroll
    for (int i = 0; i < 3 * NA; ++i) dxi[i] = rx.xi_ideal * ds1[i] + (1.0 - rx.xi_ideal) * ds0[i];
    return rx.xi_ideal * s1 + (1.0 - rx.xi_ideal) * s0;
}

// Product of the Hessians of s1 and s0 with a vector v (3*natoms): the
// reference assembles both matrices (src/calc_xi.f90:283-480); their blocks
// are those of a distance, (|R|^2 I - R R^T) / |R|^3, so the product needs
// only the differences of v across each bond and across each pair of
// fragment centres of mass (v weighted as the centre of mass weights
// positions).
template <int NA>
__device__ void xi_surface_hess_vec(const Rx &rx, const double *c, const double *v, double *h0v, double *h1v)
{
#pragma unroll
    for (int i = 0; i < 3 * NA; ++i) h0v[i] = h1v[i] = 0.0;
    auto dist_block = [](const double *R, const double *u, double sg, double *w) {
        const double rr = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
        const double rinv = 1.0 / sqrt(rr);
        const double f = sg * rinv * rinv * rinv;
        const double ru = R[0] * u[0] + R[1] * u[1] + R[2] * u[2];
#pragma unroll
        for (int d = 0; d < 3; ++d) w[d] = (rr * u[d] - R[d] * ru) * f;
    };
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
        for (int i = 0; i < CAM_MAX_BONDS; ++i) {
            const int nbond = pass == 0 ? rx.n_form : rx.n_break;
            if (i >= nbond) continue;
            const double sg = pass == 0 ? -rx.inv_nform : rx.inv_nbreak;   // forming bonds enter s1 with -|R|
            const int a = pass == 0 ? rx.form[i][0] : rx.brk[i][0], b = pass == 0 ? rx.form[i][1] : rx.brk[i][1];
            double ca[3], cb[3], va[3], vb[3], R[3], u[3], w[3], mw[3];
            gather3<NA>(c, a, ca);
            gather3<NA>(c, b, cb);
            gather3<NA>(v, a, va);
            gather3<NA>(v, b, vb);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                R[d] = ca[d] - cb[d];
                u[d] = va[d] - vb[d];
            }
            dist_block(R, u, sg, w);
#pragma unroll
            for (int d = 0; d < 3; ++d) mw[d] = -w[d];
            scatter_add3<NA>(h1v, a, w);
            scatter_add3<NA>(h1v, b, mw);
        }
    }
    double com[CAM_MAX_REAC][3], vcom[CAM_MAX_REAC][3];
#pragma unroll
    for (int k = 0; k < CAM_MAX_REAC; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double cc = 0.0, vc = 0.0;
#pragma unroll
            for (int at = 0; at < NA; ++at) {
                cc = fma(rx.fw[k][at], c[3 * at + d], cc);
                vc = fma(rx.fw[k][at], v[3 * at + d], vc);
            }
            com[k][d] = cc;
            vcom[k][d] = vc;
        }
#pragma unroll
    for (int i = 0; i < CAM_MAX_REAC; ++i)
#pragma unroll
        for (int j = i + 1; j < CAM_MAX_REAC; ++j) {
            if (j >= rx.n_reac) continue;
            double R[3], u[3], w[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                R[d] = com[j][d] - com[i][d];
                u[d] = vcom[i][d] - vcom[j][d];
            }
            dist_block(R, u, -rx.inv_terms, w);                         // s0 = r_inf - |R|
#pragma unroll
            for (int at = 0; at < NA; ++at) {
                const double wij = rx.fw[i][at] - rx.fw[j][at];
#pragma unroll
                for (int d = 0; d < 3; ++d) h0v[3 * at + d] = fma(wij, w[d], h0v[3 * at + d]);
            }
        }
}

// umbrella mode 0 (src/umbrella.f90:109-173) at centroid c for the window
// xi_ideal: xi of mode 1, s0 / (s0 - s1) (src/calc_xi.f90:204-206,
// :275-278), returned in xi1, and its gradient into dxi; the bias gradient
// every bead receives, in the reference's two parts:
// gk = k_force (xi - xi_ideal) dxi and the metric term
// gh = -(1/beta) H M^-1 dxi / (dxi^T M^-1 dxi), H the Hessian of xi, which
// the reference contracts from the full matrix (:148-170) and which is
// formed here as a Hessian-vector product.  Returns the bias energy
// 0.5 k_force (xi - xi_ideal)^2 - log(fs) / beta.
template <int NA>
__device__ double umbrella_bias(const Rx &rx, const double *c, double beta, double k_force, double xi_ideal,
                                double &xi1, double *dxi, double *gk, double *gh)
{
    constexpr int N = 3 * NA;
    double s0, s1, ds0[N], ds1[N];
    xi_surfaces<NA, true>(rx, c, s0, s1, ds0, ds1);
    const double D = s0 - s1;
    const double inv_D = 1.0 / D;
    const double inv_D2 = inv_D * inv_D;
    const double xi = s0 * inv_D;
    xi1 = xi;
#pragma unroll
    for (int i = 0; i < N; ++i) dxi[i] = (s0 * ds1[i] - s1 * ds0[i]) * inv_D2;
    const double delta = xi - xi_ideal;
#pragma unroll
    for (int i = 0; i < N; ++i) gk[i] = k_force * delta * dxi[i];
    double v[N], fs2 = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = dxi[i] * rx.inv_mass[i / 3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int at = 0; at < NA; ++at) fs2 = fs2 + dxi[3 * at + d] * v[3 * at + d];
    const double coeff1 = 2.0 * 3.1415926535897932384 * beta, coeff2 = -1.0 / beta;
    fs2 = fs2 / coeff1;
    const double vadd = 0.5 * k_force * delta * delta + coeff2 * log(sqrt(fs2));
    double h0v[N], h1v[N];
    xi_surface_hess_vec<NA>(rx, c, v, h0v, h1v);
    double ds0v = 0.0, ds1v = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        ds0v += ds0[i] * v[i];
        ds1v += ds1[i] * v[i];
    }
    const double cross = s0 * ds1v - s1 * ds0v;
    const double f = coeff2 / (coeff1 * fs2) * inv_D2 * inv_D;
#pragma unroll
    for (int i = 0; i < N; ++i)
        gh[i] = ((s0 * h1v[i] + ds0[i] * ds1v - ds1[i] * ds0v - s1 * h0v[i]) * D - 2.0 * cross * (ds0[i] - ds1[i])) * f;
    return vadd;
}

// ---------------------------------------------------------------------------
// Normal-mode step on the matrix cores (FP64 MFMA 16x16x4), nbeads 16, 32 or
// 64.  The block's p and q columns are staged in LDS as [column][bead] with
// column = 2*(tl*NC + coord) + (0: p, 1: q) and an odd stride NB + 1, so
// that the 16 columns of a tile fall in distinct banks.  Each
// wave owns one 16-row band of the transform (its A operand stays in
// registers) and the column tiles c0, c0 + CPW, ... of that band, whose
// accumulation chains it interleaves so that consecutive MFMAs are
// independent.  Pass 0 computes the modes Y = T X into ys and applies the
// exact 2x2 free ring-polymer map to the accumulators (p and q of one
// coordinate sit in neighbouring lanes, swapped with DPP); pass 1 transforms
// ys back into xs.
// Operand layout of v_mfma_f64_16x16x4f64: lane l supplies A[l%16][l/16] and
// B[l/16][l%16] and receives D[l/16 + 4r][l%16] in accumulator element r.
// ---------------------------------------------------------------------------

template <int NB>
constexpr int nm_stride() { return NB + 1; }

template <int NB, int TPB, int NC, bool CENT>
__device__ __forceinline__ void nm_mfma(double *xs, double *ys, const double *prop, const double *tfwd,
                                        const double *tinv, double *cent)
{
    constexpr int NA = NC / 3;
    constexpr int NBP = nm_stride<NB>();
    constexpr int NCOL = TPB * 2 * NC;
    constexpr int CT = (NCOL + 15) / 16;
    constexpr int RT = NB / 16;
    constexpr int NWAVE = kBlock / 64;
    static_assert(NB % 16 == 0 && NWAVE % RT == 0, "row bands must map onto whole waves");
    constexpr int CPW = NWAVE / RT;                 // waves sharing a row band
    constexpr int MAXT = (CT + CPW - 1) / CPW;      // column tiles per wave
    constexpr int PW = PropTab<NA>::W;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int li = lane & 15, lk = lane >> 4;
    const int row0 = (wave % RT) * 16, c0 = wave / RT;
#pragma unroll 1
    for (int pass = 0; pass < 2; ++pass) {
        const double *T = pass == 0 ? tfwd : tinv;
        const double *src = pass == 0 ? xs : ys;
        double *dst = pass == 0 ? ys : xs;
        double av[NB / 4];
#pragma unroll
        for (int kk = 0; kk < NB / 4; ++kk) av[kk] = T[(4 * kk + lk) * NB + row0 + li];   // T[row][k]
        dbl4 acc[MAXT];
#pragma unroll
        for (int j = 0; j < MAXT; ++j) acc[j] = dbl4{ 0.0, 0.0, 0.0, 0.0 };
        // Every wave runs MAXT tiles, so that no MFMA sits under a branch; a
        // tile past the last one reads zeros and stores nothing.
#pragma unroll
        for (int kk = 0; kk < NB / 4; ++kk)
#pragma unroll
            for (int j = 0; j < MAXT; ++j) {
                const int col = (c0 + j * CPW) * 16 + li;
                const double bv = col < NCOL ? src[col * NBP + 4 * kk + lk] : 0.0;   // X[k][col]
                acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[kk], bv, acc[j], 0, 0, 0);
            }
#pragma unroll
        for (int j = 0; j < MAXT; ++j) {
            const int col = (c0 + j * CPW) * 16 + li;
            const bool ok = col < NCOL;
            if (pass == 0) {
                // free ring-polymer map of each mode on (p, q) (src/verlet.f90:434-447)
                const bool isq = (col & 1) != 0;
                const int pair = (ok ? col : 0) >> 1;
                const int tl = pair / NC, at = (pair % NC) / 3;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const double other = quad::dpp<quad::perm(1, 0, 3, 2)>(acc[j][r]);
                    const int mode = row0 + lk + 4 * r;
                    const double *pr = prop + (tl * NB + mode) * PW;
                    const double cw = pr[0], p2 = pr[1 + 2 * at], p3 = pr[2 + 2 * at];
                    acc[j][r] = isq ? other * p3 + acc[j][r] * cw : acc[j][r] * cw + other * p2;
                }
                // The centroid of the propagated beads is mode 0 of q over sqrt(NB):
                // every column of either transform sums to sqrt(NB) times delta_k0.
                if (CENT && isq && ok && row0 == 0 && lk == 0) cent[pair] = acc[j][0] * (1.0 / kSqrt<NB>());
            }
            if (ok) {
#pragma unroll
                for (int r = 0; r < 4; ++r) dst[col * NBP + row0 + lk + 4 * r] = acc[j][r];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// The step kernel.  ENS: CAM_ENS_NVE, CAM_ENS_ANDERSEN, CAM_ENS_CHILD,
// CAM_ENS_UMBRELLA or CAM_ENS_NHC (the chain before the first and after the
// second half kick).  One step: p -= dt/2 g; free ring-polymer propagation in
// normal-mode space; [CHILD: xi of the centroid]; forces; [UMBRELLA: the
// window's bias at the centroid, added to every bead]; p -= dt/2 g;
// [ANDERSEN, UMBRELLA: resample].
//
// SPLIT: the same step without its force phase, one step per launch, for the
// split schedule (forces in force_kernel at its own register budget).  The
// launch first finishes the previous step of the run (p -= dt/2 g with that
// step's forces, Andersen), then kicks, propagates and evaluates xi; the run's
// last half kick is final_kick_kernel.
// ---------------------------------------------------------------------------
template <class S, int NB, typename T, int ENS, int WPE, bool MF, bool SPLIT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WPE))) void step_kernel(StepArgs a)
{
    constexpr int NC = S::NC, L = S::L, SL = S::SLOTS;
    constexpr int BPB = kBlock / L;          // beads per block
    constexpr int TPB = BPB / NB;            // trajectories per block
    static_assert(TPB >= 1, "a trajectory must fit in one block");
    constexpr int XS = MF ? TPB * 2 * NC * nm_stride<NB>() : TPB * NC * NB;
    constexpr int PW = PropTab<S::NA>::W;
    __shared__ __align__(16) double X[XS];
    __shared__ __align__(16) double cent[TPB * NC];
    __shared__ __align__(16) double Y[MF ? XS : 1];
    // With two waves per SIMD (WPE 2) the step keeps nothing but the state in
    // registers: the propagator table and the masses come from LDS, and q, p
    // wait in LDS while the forces are evaluated, so that the force phase has
    // the register budget of the force kernel alone.
    constexpr bool STASH = WPE >= 2;
    constexpr bool TAB = MF || STASH;
    __shared__ double prop[TAB ? TPB * NB * PW : 1];
    __shared__ double smass[TAB ? S::NA : 1];
    __shared__ double stash[STASH ? 2 * SL * kBlock : 1];
    __shared__ double wcom[ENS == CAM_ENS_CHILD ? CAM_MAX_REAC : 1][S::NA];
    // UMBRELLA: the two parts of each trajectory's bias gradient, its bead
    // energies and a flag for positions that are not finite
    constexpr bool UMB = ENS == CAM_ENS_UMBRELLA;
    constexpr bool CENT = ENS == CAM_ENS_CHILD || UMB;
    __shared__ double ubias[UMB ? TPB * 2 * NC : 1];
    __shared__ double uvb[UMB ? TPB * NB : 1];
    __shared__ int ubad[UMB ? TPB : 1];
    // NHC: the lanes' kinetic-energy parts and each trajectory's scale factor
    constexpr bool NHC = ENS == CAM_ENS_NHC;
    __shared__ double nred[NHC ? kBlock : 1];
    __shared__ double nscale[NHC ? TPB : 1];

    const int tid = threadIdx.x;
    const int r = tid % L;
    const int bib = tid / L;                 // bead in block
    // When NB does not divide the block's bead slots, the slots past the last
    // whole ring polymer repeat beads 0.. of the block's last one: they
    // compute the same values as those lanes, write the same values to LDS,
    // and nothing to memory.
    const int tl = min(bib / NB, TPB - 1), k = bib % NB;
    const bool live = bib < TPB * NB;        // not such a repeat: does the per-trajectory work
    const int traj_raw = blockIdx.x * TPB + tl;
    const bool act = live && traj_raw < a.ntraj;
    const int traj = traj_raw < a.ntraj ? traj_raw : a.ntraj - 1;
    double *xs = X + tl * NC * NB;

    int cc[SL];
    double m[SL], q[SL], p[SL], g[SL];
    const size_t base = (size_t)traj * NC * NB;
#pragma unroll
    for (int s = 0; s < SL; ++s) {
        cc[s] = S::coord(s, r);
        m[s] = a.mass[cc[s] / 3];
        const bool ok = S::valid(s, r);
        q[s] = ok ? a.q[base + (size_t)cc[s] * NB + k] : 0.0;
        p[s] = ok ? a.p[base + (size_t)cc[s] * NB + k] : 0.0;
        g[s] = ok ? a.g[base + (size_t)cc[s] * NB + k] : 0.0;
    }
    // free ring-polymer propagator of mode k (src/verlet.f90:405-432)
    const double beta = a.beta[traj];
    double p1 = 1.0, p2[SL], p3[SL];
    if constexpr (TAB) {
        // coefficients come from the LDS table inside the step loop
    } else if (NB > 1 && k != 0) {
        const int kk = k <= NB / 2 ? k : NB - k;
        const double wk = (2.0 / (beta / NB)) * sin(kk * (3.141592653589793 / NB));
        const double wt = wk * a.dt;
        const double cw = cos(wt), sw = sin(wt);
        p1 = cw;
#pragma unroll
        for (int s = 0; s < SL; ++s) {
            const double wm = wk * m[s];
            p2[s] = -wm * sw;
            p3[s] = sw / wm;
        }
    } else {
#pragma unroll
        for (int s = 0; s < SL; ++s) {
            p2[s] = 0.0;
            p3[s] = a.dt / m[s];
        }
    }
    const double hdt = 0.5 * a.dt;
    double vb = 0.0;
    double usum = 0.0, usq = 0.0;            // UMBRELLA: this launch's sums of xi, xi^2
    bool ucheck = false;                     // UMBRELLA: the window check failed in this launch
    // NHC: the chain of this lane's trajectory lives in the registers of its
    // bead-0 lane for the launch; masses and kT as mdinit and nhc set them
    // (src/mdinit.f90:139-146; src/nhc.f90:53, a default-REAL quotient)
    double vnh[4] = { 0.0, 0.0, 0.0, 0.0 }, qnh[4] = { 0.0, 0.0, 0.0, 0.0 }, ekt_chain = 0.0;
    if constexpr (NHC) {
        const double kelvin = 1.0 / (beta * 0.316679e-5);
        const double qterm = 0.316679e-5 * kelvin * a.nose_q * a.nose_q;
        qnh[0] = (double)NC * qterm;
        qnh[1] = qnh[2] = qnh[3] = qterm;
        ekt_chain = (double)(1.380649E-23f / 4.3597447E-18f) * kelvin;
        if (k == 0 && r == 0)
#pragma unroll
            for (int j = 0; j < 4; ++j) vnh[j] = a.nhc[4 * (size_t)traj + j];
    }
    // one application of the chain to this block's trajectories
    auto nhc_apply = [&]() {
        double part = 0.0;
#pragma unroll
        for (int s = 0; s < SL; ++s)
            if (S::valid(s, r)) part += p[s] * p[s] / (2.0 * (TAB ? smass[cc[s] / 3] : m[s]));
        nred[tid] = part;
        __syncthreads();
        if (live && k == 0 && r == 0) {
            double ek = 0.0;
            for (int i = 0; i < NB * L; ++i) ek += nred[tl * NB * L + i];
            nscale[tl] = nhc_chain(vnh, qnh, ek / NB / NB, ekt_chain, NC, a.dt);
        }
        __syncthreads();
        const double f = nscale[tl];
#pragma unroll
        for (int s = 0; s < SL; ++s) p[s] = f * p[s];
    };
    if constexpr (ENS == CAM_ENS_CHILD) {
        // fragment weights of the centre-of-mass sums (calc_com, src/calc_com.f90:36-59)
        if (threadIdx.x < CAM_MAX_REAC * S::NA) {
            const int kf = threadIdx.x / S::NA, at = threadIdx.x % S::NA;
            double w = 0.0;
            if (kf < a.rx.n_reac)
                for (int i = 0; i < a.rx.reac_n[kf]; ++i)
                    if (a.rx.reac_at[kf][i] == at) w = a.mass[at] / a.rx.reac_mass[kf];
            wcom[kf][at] = w;
        }
    }
    if constexpr (TAB) {
        if (threadIdx.x < S::NA) smass[threadIdx.x] = a.mass[threadIdx.x];
        // per-trajectory mode table (src/verlet.f90:413-432), as PropTab lays it out
        for (int i = threadIdx.x; i < TPB * NB; i += kBlock) {
            const int t = i / NB, kk0 = i % NB;
            const int tr = min(blockIdx.x * TPB + t, a.ntraj - 1);
            const int kk = kk0 <= NB / 2 ? kk0 : NB - kk0;
            const double bt = a.beta[tr];
            const double wk = (2.0 / (bt / NB)) * sin(kk * (3.141592653589793 / NB));
            const double wt = wk * a.dt;
            const double cw = cos(wt), sw = sin(wt);
            double *pr = prop + i * PW;
            pr[0] = kk0 == 0 ? 1.0 : cw;
#pragma unroll
            for (int at = 0; at < S::NA; ++at) {
                const double m = a.mass[at];
                if (kk0 == 0) {
                    pr[1 + 2 * at] = 0.0;
                    pr[2 + 2 * at] = a.dt / m;
                } else {
                    const double wm = wk * m;
                    pr[1 + 2 * at] = -wm * sw;
                    pr[2 + 2 * at] = sw / wm;
                }
            }
        }
    }

    if constexpr (UMB) {
        for (int i = tid; i < TPB; i += kBlock) ubad[i] = 0;
    }
    if constexpr (SPLIT) {
        if (a.step_off > 0) {
            // second half kick of the previous step (src/verlet.f90:1060-1062) and its Andersen draw
#pragma unroll
            for (int s = 0; s < SL; ++s) p[s] = p[s] - hdt * g[s];
            if constexpr (ENS == CAM_ENS_ANDERSEN) {
                const uint64_t istep = *a.step_base + (uint64_t)a.step_off;
                if (istep % (uint64_t)a.andersen_step == 0) {
                    __syncthreads();      // the mass table is written before the first step
#pragma unroll
                    for (int s = 0; s < SL; ++s) {
                        const double sd = sqrt((TAB ? smass[cc[s] / 3] : m[s]) / (beta / NB));
                        p[s] = andersen_normal(a.seed, (uint32_t)traj, istep, cc[s], k) * sd;
                    }
                }
            }
        }
    }
    for (int it = 0; it < a.nsteps; ++it) {
        if constexpr (NHC) {
            if constexpr (TAB) __syncthreads();       // the mass table is written before the first step
            nhc_apply();                              // src/verlet.f90:185-206
        }
        // half kick (src/verlet.f90:216-218)
#pragma unroll
        for (int s = 0; s < SL; ++s) p[s] = p[s] - hdt * g[s];

        if constexpr (NB == 1) {
            if constexpr (TAB) {
                __syncthreads();          // the table is written before the first step
#pragma unroll
                for (int s = 0; s < SL; ++s) q[s] = p[s] * (a.dt / smass[cc[s] / 3]) + q[s];
            } else {
#pragma unroll
                for (int s = 0; s < SL; ++s) q[s] = p[s] * p3[s] + q[s];
            }
        } else if constexpr (MF) {
            double *xm = X + tl * 2 * NC * nm_stride<NB>();
            __syncthreads();
#pragma unroll
            for (int s = 0; s < SL; ++s)
                if (S::valid(s, r)) {
                    xm[(2 * cc[s]) * nm_stride<NB>() + k] = p[s];
                    xm[(2 * cc[s] + 1) * nm_stride<NB>() + k] = q[s];
                }
            __syncthreads();
            nm_mfma<NB, TPB, NC, CENT>(X, Y, prop, a.tfwd, a.tinv, cent);
#pragma unroll
            for (int s = 0; s < SL; ++s) {
                p[s] = xm[(2 * cc[s]) * nm_stride<NB>() + k];
                q[s] = xm[(2 * cc[s] + 1) * nm_stride<NB>() + k];
            }
        } else {
            // bead -> normal modes, propagate, back (src/verlet.f90:377-463)
            double yp[SL], yq[SL];
#pragma unroll
            for (int pass = 0; pass < 2; ++pass) {
                __syncthreads();
#pragma unroll
                for (int s = 0; s < SL; ++s)
                    if (S::valid(s, r)) xs[cc[s] * NB + k] = pass == 0 ? p[s] : q[s];
                __syncthreads();
                double y[SL];
#pragma unroll
                for (int s = 0; s < SL; ++s) y[s] = 0.0;
#pragma unroll 1
                for (int j = 0; j < NB; j += 2) {
                    const double t0 = a.tfwd[j * NB + k], t1 = a.tfwd[(j + 1) * NB + k];
#pragma unroll
                    for (int s = 0; s < SL; ++s) {
                        const double2 x2 = *reinterpret_cast<const double2 *>(&xs[cc[s] * NB + j]);
                        y[s] = fma(t0, x2.x, y[s]);
                        y[s] = fma(t1, x2.y, y[s]);
                    }
                }
#pragma unroll
                for (int s = 0; s < SL; ++s) {
                    if (pass == 0) yp[s] = y[s];
                    else yq[s] = y[s];
                }
            }
            if constexpr (TAB) {
                // the same coefficients, from the table (mode k = this lane's bead index)
                const double *pr = prop + (tl * NB + k) * PW;
                const double cw = pr[0];
#pragma unroll
                for (int s = 0; s < SL; ++s) {
                    const int at = cc[s] / 3;
                    const double c2 = pr[1 + 2 * at], c3 = pr[2 + 2 * at];
                    const double pn = yp[s] * cw + yq[s] * c2;
                    yq[s] = yp[s] * c3 + yq[s] * cw;
                    yp[s] = pn;
                }
            } else {
#pragma unroll
                for (int s = 0; s < SL; ++s) {
                    const double pn = yp[s] * p1 + yq[s] * p2[s];
                    yq[s] = yp[s] * p3[s] + yq[s] * p1;
                    yp[s] = pn;
                }
            }
            if constexpr (CENT) {
                // centroid of the propagated beads = mode 0 of q / sqrt(NB)
                if (live && k == 0) {
#pragma unroll
                    for (int s = 0; s < SL; ++s)
                        if (S::valid(s, r)) cent[tl * NC + cc[s]] = yq[s] * (1.0 / kSqrt<NB>());
                }
            }
#pragma unroll
            for (int pass = 0; pass < 2; ++pass) {
                __syncthreads();
#pragma unroll
                for (int s = 0; s < SL; ++s)
                    if (S::valid(s, r)) xs[cc[s] * NB + k] = pass == 0 ? yp[s] : yq[s];
                __syncthreads();
                double y[SL];
#pragma unroll
                for (int s = 0; s < SL; ++s) y[s] = 0.0;
#pragma unroll 1
                for (int j = 0; j < NB; j += 2) {
                    const double t0 = a.tinv[j * NB + k], t1 = a.tinv[(j + 1) * NB + k];
#pragma unroll
                    for (int s = 0; s < SL; ++s) {
                        const double2 x2 = *reinterpret_cast<const double2 *>(&xs[cc[s] * NB + j]);
                        y[s] = fma(t0, x2.x, y[s]);
                        y[s] = fma(t1, x2.y, y[s]);
                    }
                }
#pragma unroll
                for (int s = 0; s < SL; ++s) {
                    if (pass == 0) p[s] = y[s];
                    else q[s] = y[s];
                }
            }
        }

        if constexpr (ENS == CAM_ENS_CHILD) {
            // centroid (src/get_centroid.f90:66-85) and xi of the child step
            // (umbrella mode 1 -> calc_xi mode 2, src/verlet.f90:1049-1050).
            // With beads, the normal-mode step above left the centroid in cent
            // behind a barrier; one bead is its own centroid.
            if constexpr (NB == 1) {
                __syncthreads();
#pragma unroll
                for (int s = 0; s < SL; ++s)
                    if (S::valid(s, r)) cent[tl * NC + cc[s]] = q[s];
                __syncthreads();
            }
            if constexpr (SPLIT) {
                if (k == 0 && r == 0 && act) {
                    const double xi = calc_xi_value<S::NA>(a.rx, wcom, cent + tl * NC);
                    a.xi_flag[(size_t)(a.step_off + it) * a.ntraj + traj] = xi > 0.0 ? 1 : 0;
                }
            } else {
                // xi of these centroids is evaluated by xi_flag_kernel after the
                // launch, at full occupancy, instead of by one lane while its SIMD waits
                for (int i = tid; i < TPB * NC; i += kBlock) {
                    const int tr = blockIdx.x * TPB + i / NC;
                    if (tr < a.ntraj) a.cbuf[((size_t)it * a.ntraj + tr) * NC + i % NC] = cent[i];
                }
            }
        }

        if constexpr (UMB && NB == 1) {
            __syncthreads();
#pragma unroll
            for (int s = 0; s < SL; ++s)
                if (S::valid(s, r)) cent[tl * NC + cc[s]] = q[s];
        }
        if constexpr (SPLIT) continue;     // one step per launch; forces come next
        // forces of every bead (src/verlet.f90:772-777)
        {
            double x[SL];
#pragma unroll
            for (int s = 0; s < SL; ++s) x[s] = q[s];
            if constexpr (STASH) {
#pragma unroll
                for (int s = 0; s < SL; ++s) {
                    stash[s * kBlock + tid] = q[s];
                    stash[(SL + s) * kBlock + tid] = p[s];
                }
            }
            S::template egrad<T>(x, vb, g);
            if constexpr (STASH) {
#pragma unroll
                for (int s = 0; s < SL; ++s) {
                    q[s] = stash[s * kBlock + tid];
                    p[s] = stash[(SL + s) * kBlock + tid];
                }
            }
        }
        if constexpr (UMB) {
            // umbrella mode 0 at the centroid of this step's positions
            // (src/verlet.f90:1046, src/umbrella.f90:109-173): one lane per
            // trajectory evaluates it; every bead adds it, the force-constant
            // part first.  That lane also sums xi and runs the window check
            // (rpmd_check traj_type 2/3, src/rpmd_check.f90:74-112).
            bool qbad = false;
#pragma unroll
            for (int s = 0; s < SL; ++s) qbad = qbad || (S::valid(s, r) && (q[s] != q[s] || q[s] > DBL_MAX));
            if (live && r == 0) uvb[tl * NB + k] = vb;
            if (live && qbad) ubad[tl] = 1;
            __syncthreads();
            if (live && k == 0 && r == 0) {
                double dxi[NC], gk[NC], gh[NC], xi1;
                const double xw = a.uwin[2 * traj], kw = a.uwin[2 * traj + 1];
                (void)umbrella_bias<S::NA>(a.rx, cent + tl * NC, beta, kw, xw, xi1, dxi, gk, gh);
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    ubias[(2 * tl) * NC + i] = gk[i];
                    ubias[(2 * tl + 1) * NC + i] = gh[i];
                }
                usum += xi1;
                usq += xi1 * xi1;
                // slots past the last trajectory evaluate a copy of it: not counted
                if (act) count_xi(a.uhist, a.hbins, a.hmin, a.hinv, traj, xi1);
                const uint64_t istep = a.step_base[1] + (uint64_t)(a.step_off + it) + 1;
                if (a.check_after >= 0 && istep > (uint64_t)a.check_after) {
                    double ep = 0.0;
                    for (int kk = 0; kk < NB; ++kk) ep = ep + uvb[tl * NB + kk];
                    const bool bad = ubad[tl] != 0 || ep != ep || ep > DBL_MAX || ep > a.epot_max ||
                                     fabs(xi1 - xw) > a.xi_tol;
                    if (bad) ucheck = true;
                }
                ubad[tl] = 0;
            }
            __syncthreads();
#pragma unroll
            for (int s = 0; s < SL; ++s)
                if (S::valid(s, r)) g[s] = (g[s] + ubias[(2 * tl) * NC + cc[s]]) + ubias[(2 * tl + 1) * NC + cc[s]];
        }
        // second half kick (src/verlet.f90:1060-1062)
#pragma unroll
        for (int s = 0; s < SL; ++s) p[s] = p[s] - hdt * g[s];

        if constexpr (NHC) nhc_apply();             // src/verlet.f90:1086-1106
        if constexpr (UMB) {
            // Andersen every andersen_step steps counted from the segment's
            // start, the draw keyed by the steps since the engine's creation
            if (a.andersen_step > 0) {
                const uint64_t istep = a.step_base[1] + (uint64_t)(a.step_off + it) + 1;
                const uint64_t key = a.step_base[0] + (uint64_t)(a.step_off + it) + 1;
                if (istep % (uint64_t)a.andersen_step == 0) {
#pragma unroll
                    for (int s = 0; s < SL; ++s) {
                        const double sd = sqrt((TAB ? smass[cc[s] / 3] : m[s]) / (beta / NB));
                        p[s] = andersen_normal(a.seed, (uint32_t)traj, key, cc[s], k) * sd;
                    }
                }
            }
        }
        if constexpr (ENS == CAM_ENS_ANDERSEN) {
            // Andersen: all bead momenta redrawn every andersen_step steps
            // (src/verlet.f90:1159-1165, src/andersen.f90:37-74)
            const uint64_t istep = *a.step_base + (uint64_t)(a.step_off + it) + 1;
            if (istep % (uint64_t)a.andersen_step == 0) {
#pragma unroll
                for (int s = 0; s < SL; ++s) {
                    const double sd = sqrt((TAB ? smass[cc[s] / 3] : m[s]) / (beta / NB));
                    p[s] = andersen_normal(a.seed, (uint32_t)traj, istep, cc[s], k) * sd;
                }
            }
        }
    }
    if (act) {
#pragma unroll
        for (int s = 0; s < SL; ++s)
            if (S::valid(s, r)) {
                const size_t o = base + (size_t)cc[s] * NB + k;
                a.q[o] = q[s];
                a.p[o] = p[s];
                if (!SPLIT) a.g[o] = g[s];
            }
        if (r == 0 && !SPLIT) a.vbead[(size_t)traj * NB + k] = vb;
        if constexpr (NHC) {
            if (k == 0 && r == 0)
#pragma unroll
                for (int j = 0; j < 4; ++j) a.nhc[4 * (size_t)traj + j] = vnh[j];
        }
        if constexpr (UMB) {
            if (k == 0 && r == 0) {
                a.ustat[2 * traj] = a.ustat[2 * traj] + usum;
                a.ustat[2 * traj + 1] = a.ustat[2 * traj + 1] + usq;
                if (ucheck) a.pfail[traj] = a.pfail[traj] | kPfCheck;
            }
        }
    }
}

// Last half kick of a split-schedule run (and its Andersen draw), over the
// engine layout [traj][coord][bead].
__global__ void final_kick_kernel(StepArgs a, int natoms, int nb)
{
    const size_t nc = (size_t)3 * natoms;
    const size_t n = (size_t)a.ntraj * nc * nb;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = (int)(i % nb);
    const int c = (int)((i / nb) % nc);
    const int traj = (int)(i / (nc * nb));
    double p = a.p[i] - 0.5 * a.dt * a.g[i];
    if (a.andersen_step > 0) {
        const uint64_t istep = *a.step_base + (uint64_t)a.nsteps;
        if (istep % (uint64_t)a.andersen_step == 0) {
            const double sd = sqrt(a.mass[c / 3] / (a.beta[traj] / nb));
            p = andersen_normal(a.seed, (uint32_t)traj, istep, c, k) * sd;
        }
    }
    a.p[i] = p;
}

// xi flags of the child steps of one fused launch: one thread per (step,
// trajectory) on the centroids the step kernel wrote.
template <int NA>
__global__ __launch_bounds__(256) void xi_flag_kernel(StepArgs a)
{
    __shared__ double wcom[CAM_MAX_REAC][NA];
    if (threadIdx.x < CAM_MAX_REAC * NA) {
        const int kf = threadIdx.x / NA, at = threadIdx.x % NA;
        double w = 0.0;
        if (kf < a.rx.n_reac)
            for (int i = 0; i < a.rx.reac_n[kf]; ++i)
                if (a.rx.reac_at[kf][i] == at) w = a.mass[at] / a.rx.reac_mass[kf];
        wcom[kf][at] = w;
    }
    __syncthreads();
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = a.tend - a.tbegin;
    if (i >= (size_t)a.nsteps * n) return;
    const int it = (int)(i / n), traj = a.tbegin + (int)(i % n);
    double c[3 * NA];
#pragma unroll
    for (int j = 0; j < 3 * NA; ++j) c[j] = a.cbuf[((size_t)it * a.ntraj + traj) * 3 * NA + j];
    const double xi = calc_xi_value<NA>(a.rx, wcom, c);
    a.xi_flag[(size_t)(a.step_off + it) * a.ntraj + traj] = xi > 0.0 ? 1 : 0;
}

// Forces and bead energies of the current positions (mdinit, src/mdinit.f90:79-83;
// also the force phase of the split schedule).  WPE: waves per SIMD the
// register budget is sized for.
template <class S, int NB, typename T, int WPE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WPE))) void force_kernel(
    const double *__restrict__ qa, double *__restrict__ ga, double *__restrict__ vbead, int ntraj)
{
    constexpr int NC = S::NC, L = S::L, SL = S::SLOTS, BPB = kBlock / L, TPB = BPB / NB;
    // slots past the last whole ring polymer repeat beads of the last one and write nothing
    const int tid = threadIdx.x, r = tid % L, bib = tid / L, tl = min(bib / NB, TPB - 1), k = bib % NB;
    const int traj_raw = blockIdx.x * TPB + tl;
    const bool act = bib < TPB * NB && traj_raw < ntraj;
    const int traj = traj_raw < ntraj ? traj_raw : ntraj - 1;
    const size_t base = (size_t)traj * NC * NB;
    double x[SL], g[SL], vb;
#pragma unroll
    for (int s = 0; s < SL; ++s) x[s] = S::valid(s, r) ? qa[base + (size_t)S::coord(s, r) * NB + k] : 0.0;
    S::template egrad<T>(x, vb, g);
    if (act) {
#pragma unroll
        for (int s = 0; s < SL; ++s)
            if (S::valid(s, r)) ga[base + (size_t)S::coord(s, r) * NB + k] = g[s];
        if (r == 0) vbead[(size_t)traj * NB + k] = vb;
    }
}
}  // namespace

namespace {

template <class S, int NB, typename T, int WPE, bool SPLIT = false>
step_fn pick_step_w(int ens, bool mfma)
{
    constexpr bool MF = mfma_bead_count(NB);
    // umbrella windows and the Nose-Hoover chain: the fused step only, with
    // the vector-ALU transform
    if constexpr (!SPLIT && WPE == 1) {
        if (ens == CAM_ENS_UMBRELLA) return step_kernel<S, NB, T, CAM_ENS_UMBRELLA, 1, false, false>;
        if (ens == CAM_ENS_NHC) return step_kernel<S, NB, T, CAM_ENS_NHC, 1, false, false>;
    }
    if (mfma && MF) {
        if (ens == CAM_ENS_NVE) return step_kernel<S, NB, T, CAM_ENS_NVE, WPE, MF, SPLIT>;
        if (ens == CAM_ENS_ANDERSEN) return step_kernel<S, NB, T, CAM_ENS_ANDERSEN, WPE, MF, SPLIT>;
        return step_kernel<S, NB, T, CAM_ENS_CHILD, WPE, MF, SPLIT>;
    }
    if (ens == CAM_ENS_NVE) return step_kernel<S, NB, T, CAM_ENS_NVE, WPE, false, SPLIT>;
    if (ens == CAM_ENS_ANDERSEN) return step_kernel<S, NB, T, CAM_ENS_ANDERSEN, WPE, false, SPLIT>;
    return step_kernel<S, NB, T, CAM_ENS_CHILD, WPE, false, SPLIT>;
}

// wpe: minimum waves per SIMD the compiler is asked to fit (1 or 2).  Only
// the FP64 CH5 step offers 2 (here and in the folded kernel): it needs more
// than 256 registers at 1; which is faster is measured (cam_engine_autotune).
template <class S, int NB, typename T>
step_fn pick_step(int ens, int wpe, bool mfma)
{
    if constexpr (S::L == 4 && std::is_same<T, double>::value)
        if (wpe == 2 && ens != CAM_ENS_UMBRELLA && ens != CAM_ENS_NHC) return pick_step_w<S, NB, T, 2>(ens, mfma);
    return pick_step_w<S, NB, T, 1>(ens, mfma);
}

// The force kernel of the split schedule gets two waves per SIMD for the FP64
// CH5 surface (it then spills a few bytes), the default budget otherwise.
template <class S, int NB, typename T>
force_fn pick_force_split()
{
    if constexpr (S::L == 4 && std::is_same<T, double>::value) return force_kernel<S, NB, T, 2>;
    else return force_kernel<S, NB, T, 1>;
}

template <class S, typename T>
// ens: the fused step's ensemble; ens_nm: that of the split schedule's
// normal-mode kernel.
bool pick_nb(int nb, int ens, int ens_nm, int wpe, bool mfma, step_fn &st, force_fn &fo, step_fn &nm, force_fn &fs)
{
    switch (nb) {
#define CAM_NB_CASE(N)                                                   \
    case N:                                                              \
        st = pick_step<S, N, T>(ens, wpe, mfma);                         \
        fo = force_kernel<S, N, T, 1>;                                   \
        nm = pick_step_w<S, N, double, 1, true>(ens_nm, mfma);           \
        fs = pick_force_split<S, N, T>();                                \
        return true;
        CAM_NB_CASE(1)
        CAM_NB_CASE(2)
        CAM_NB_CASE(4)
        CAM_NB_CASE(6)
        CAM_NB_CASE(8)
        CAM_NB_CASE(12)
        CAM_NB_CASE(16)
        CAM_NB_CASE(24)
        CAM_NB_CASE(32)
        CAM_NB_CASE(48)
        CAM_NB_CASE(64)
#undef CAM_NB_CASE
    default:
        return false;
    }
}
}  // namespace
}  // namespace cam
