// Definitions shared by the step kernels of the trajectory-batch engine
// (engine.hip) and the folded step kernel (fold.hip): the launch block size,
// how a surface spreads the coordinates of one bead over its lanes, the
// kernel arguments and the layout of the per-block propagator table.
#pragma once
#include <cstdint>

#include "common.hpp"
#include "pes_brh2.hpp"
#include "pes_ch4h.hpp"
#include "pes_h3.hpp"
#include "pes_o3.hpp"
#include "pes_oh3.hpp"
#include "nhc.hpp"

namespace cam {

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// Surface policies: how the coordinates of one bead are spread over its L lanes.
// ---------------------------------------------------------------------------
// H3: one lane per bead; T = float evaluates the surface's asymmetric and
// compact correction terms in binary32 (mixed precision).  Every translation
// unit that instantiates a step kernel with it is compiled without
// machine-level loop-invariant code motion: with that pass, those kernels
// (inlined surface in a multi-step loop, hundreds of hoisted 64-bit constants
// spilled from SGPRs to VGPR lanes) returned NaN forces in 12 of 26
// configurations, and without it all 26 are within 2e-7 of the oracle.
struct SurfH3 {
    static constexpr int NA = 3, NC = 9, L = 1, SLOTS = 9;
    __device__ static int coord(int slot, int) { return slot; }
    __device__ static bool valid(int, int) { return true; }
    template <typename T>
    __device__ static void egrad(const double *own, double &v, double *g) { h3::egrad_bead<T>(own, v, g); }
};

// OH + H2: one lane per bead, FP64 only (the engine rejects mixed precision).
struct SurfOH3 {
    static constexpr int NA = 4, NC = 12, L = 1, SLOTS = 12;
    __device__ static int coord(int slot, int) { return slot; }
    __device__ static bool valid(int, int) { return true; }
    template <typename T>
    __device__ static void egrad(const double *own, double &v, double *g) { oh3::egrad_bead(own, v, g); }
};

// BrH2 DIM-3C: one lane per bead, FP64 only (the engine rejects mixed precision).
struct SurfBrH2 {
    static constexpr int NA = 3, NC = 9, L = 1, SLOTS = 9;
    __device__ static int coord(int slot, int) { return slot; }
    __device__ static bool valid(int, int) { return true; }
    template <typename T>
    __device__ static void egrad(const double *own, double &v, double *g) { brh2::egrad_bead(own, v, g); }
};

// O + O2 (Varga-Paukku-Truhlar): one lane per bead, FP64 only.
struct SurfO3 {
    static constexpr int NA = 3, NC = 9, L = 1, SLOTS = 9;
    __device__ static int coord(int slot, int) { return slot; }
    __device__ static bool valid(int, int) { return true; }
    template <typename T>
    __device__ static void egrad(const double *own, double &v, double *g) { o3::egrad_bead(own, v, g); }
};

struct SurfCH4H {
    static constexpr int NA = 6, NC = 18, L = 4, SLOTS = 6;
    // slot 0..2: methane H number r; slot 3..5: C on lane 0, Hb on lane 1
    __device__ static int coord(int slot, int r)
    {
        return slot < 3 ? 3 * ch4h::hidx(r) + slot : 3 * (r == 0 ? ch4h::kC : ch4h::kB) + slot - 3;
    }
    __device__ static bool valid(int slot, int r) { return slot < 3 || r < 2; }
    template <typename T>
    __device__ static void egrad(const double *own, double &v, double *g) { ch4h::egrad_quad<T>(own, v, g); }
};

struct Rx {                                  // reaction coordinate, 0-based indices
    int n_form, n_break, n_reac;
    int form[CAM_MAX_BONDS][2], brk[CAM_MAX_BONDS][2];
    double form_ref[CAM_MAX_BONDS], break_ref[CAM_MAX_BONDS];
    int reac_n[CAM_MAX_REAC], reac_at[CAM_MAX_REAC][CAM_MAX_ATOMS];
    double reac_mass[CAM_MAX_REAC];
    double r_inf, xi_ideal;
    // derived on the host: fw[k][at] = mass[at] / reac_mass[k] for the atoms
    // of reactant k (0 for the others), reciprocals of the masses and counts
    double fw[CAM_MAX_REAC][CAM_MAX_ATOMS];
    double inv_mass[CAM_MAX_ATOMS];
    double inv_nform, inv_nbreak, inv_terms;
};

// UMBRELLA: one sample xi of trajectory traj into its histogram h (hbins bins
// per trajectory): bin j = nint((xi - hmin) hinv), 1 <= j <= hbins, the bin
// the coordinate hmin + j / hinv stands for (src/wham.f90:104-106 gives bin j
// that coordinate); other samples are not counted.  Only the lane that owns
// the trajectory's xi counts it, but the add is atomic so that nothing about
// the order of lanes matters.
__device__ inline void count_xi(unsigned *h, int hbins, double hmin, double hinv, int traj, double xi)
{
    if (!h) return;
    const double j = rint((xi - hmin) * hinv);
    if (j >= 1.0 && j <= (double)hbins) atomicAdd(h + (size_t)traj * hbins + ((int)j - 1), 1u);
}

struct StepArgs {
    double *q, *p, *g, *vbead;
    const double *beta;          // [ntraj]
    const double *tfwd, *tinv;   // [nbeads][nbeads], element (j,k) = T[k][j]
    const double *tfold;         // folded reference transform [nbeads/2][nbeads/2] (fold.hip)
    uint8_t *xi_flag;            // CHILD: [nsteps_run][ntraj]
    double *cbuf;                // CHILD, fused: centroids of this launch's steps [step][traj][3*natoms]
    double *pdxi;                // PARENT: constraint direction of each trajectory [traj][3*natoms]
    int *pfail;                  // PARENT, UMBRELLA: trajectory flags
    double *ustat;               // UMBRELLA: running sums of xi and xi^2 [traj][2]
    const double *uwin;          // UMBRELLA: (xi_ideal, k_force) of each trajectory's window
    double epot_max, xi_tol;     // PARENT, UMBRELLA: trajectory check
    int check_after;             //   armed after this many steps (never if < 0)
    double *nhc;                 // NHC: thermostat velocities [traj][4]
    unsigned *uhist;             // UMBRELLA: histograms of xi [traj][hbins], or null
    int hbins;
    double hmin, hinv;           //   bin j (1-based) holds the xi with nint((xi - hmin) hinv) = j
    double nose_q;               // NHC: damping time (atomic time units)
    const uint64_t *step_base;   // device counter of completed steps (Andersen draws)
    int ntraj, nsteps, step_off; // steps in this launch, offset of its first step in the run
    int tbegin, tend;            // trajectories of this launch (folded kernels and xi_flag_kernel;
                                 // the other kernels take all ntraj)
    int andersen_step;
    uint64_t seed;
    double dt;
    double mass[CAM_MAX_ATOMS];
    Rx rx;
};

// ---------------------------------------------------------------------------
// Propagator table of a block, per (trajectory, mode): cos(w_k dt) and, per
// atom a, the two off-diagonal elements of the exact free ring-polymer map
// (src/verlet.f90:413-432):
//   p' = cw p + p2_a q,   q' = p3_a p + cw q,
//   p2_a = -w_k m_a sin(w_k dt),  p3_a = sin(w_k dt) / (w_k m_a)  (mode 0: 0, dt/m_a).
// ---------------------------------------------------------------------------
template <int NA>
struct PropTab {
    static constexpr int W = 1 + 2 * NA;      // doubles per (trajectory, mode)
};

// FP64 MFMA accumulator of v_mfma_f64_16x16x4f64
typedef double dbl4 __attribute__((ext_vector_type(4)));

// sqrt(NB), correctly rounded, for the bead counts the step kernels are built for
template <int NB>
constexpr double kSqrt()
{
    static_assert(NB == 1 || NB == 2 || NB == 4 || NB == 6 || NB == 8 || NB == 12 || NB == 16 || NB == 24 ||
                      NB == 32 || NB == 48 || NB == 64,
                  "no sqrt for this bead count");
    return NB == 1 ? 1.0 : NB == 2 ? 1.4142135623730951 : NB == 4 ? 2.0 : NB == 6 ? 2.449489742783178
         : NB == 8 ? 2.8284271247461903 : NB == 12 ? 3.4641016151377544 : NB == 16 ? 4.0
         : NB == 24 ? 4.898979485566356 : NB == 32 ? 5.656854249492381 : NB == 48 ? 6.928203230275509 : 8.0;
}

// Bead counts with step kernels: the powers of two to 64 and 6, 12, 24, 48
// (even, so that the vector-ALU transform reads bead pairs).  The MFMA
// transform needs NB a multiple of 16 whose 16-row bands map onto the four
// waves of a block (16, 32, 64).
constexpr bool bead_count_supported(int nb)
{
    return nb == 1 || nb == 2 || nb == 4 || nb == 6 || nb == 8 || nb == 12 || nb == 16 || nb == 24 || nb == 32 ||
           nb == 48 || nb == 64;
}
constexpr bool mfma_bead_count(int nb) { return nb >= 16 && nb % 16 == 0 && 4 % (nb / 16) == 0; }

typedef void (*step_fn)(StepArgs);
typedef void (*force_fn)(const double *, double *, double *, int);

// The engine's kernels for one surface, precision and bead count: the fused
// step (ensemble ens), the initial-force kernel, and the split schedule's
// normal-mode step (ensemble ens_nm) and force kernel; false if the bead count
// has none.  wpe: waves per SIMD asked of the compiler (2 only for the FP64
// CH5 step); mfma: the MFMA normal-mode transform where the bead count has
// one.  Built in engine_h3.hip, engine_ch4h.hip and engine_fp64.hip (OH3,
// BrH2, O3: FP64 only).
bool pick_kernels_h3(int precision, int nb, int ens, int ens_nm, int wpe, bool mfma, step_fn &st, force_fn &fo,
                     step_fn &nm, force_fn &fs);
bool pick_kernels_ch4h(int precision, int nb, int ens, int ens_nm, int wpe, bool mfma, step_fn &st, force_fn &fo,
                       step_fn &nm, force_fn &fs);
bool pick_kernels_fp64(int surface, int nb, int ens, int ens_nm, int wpe, bool mfma, step_fn &st, force_fn &fo,
                       step_fn &nm, force_fn &fs);

// The folded step kernel (fold.hip) of a configuration, or nullptr if it has
// none; *tpb receives its trajectories per block.  nbuf: LDS column buffers
// of the one-lane-per-bead surfaces' kernels (1: more blocks per CU; 2: one
// block per CU, for engines that share the chip with other engines); the CH5
// kernels take 2 whatever is asked.
step_fn fold_step_kernel_for(int surface, int precision, int nbeads, int ensemble, int wpe, int nbuf, int *tpb);

// Folded reference transform M[r][c] = w_c cos(2 pi r c / nbeads) / sqrt(nbeads),
// r, c < nbeads/2, w_0 = 1 and w_c = 2 otherwise, row-major.
void fold_matrix(int nbeads, double *m);

}  // namespace cam
