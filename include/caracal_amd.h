/*
 * caracal_amd.h -- C ABI of the MI355X ring-polymer MD engine.
 *
 * Plain C types only (pointers, sizes, POD structs); no HIP or torch types.
 * Every entry point returns an int status (CAM_OK == 0) and never aborts;
 * cam_last_error() returns the message of the last failure on the calling
 * thread.  Arrays use the reference's Fortran layout: a position block
 * q(3,natoms,nbeads) is column-major, so the Cartesian component runs fastest,
 * then the atom, then the bead; a batch stacks such blocks trajectory by
 * trajectory.  Units are atomic units throughout (bohr, hartree, electron
 * masses, atomic time), as in the reference.
 *
 * Two families of entry points:
 *  (1) drop-in analytic surfaces, the bead-batched plugin contract of the
 *      reference (src/egrad_h3.f:29-35, src/egrad_ch4h.f:74-79);
 *  (2) a trajectory-batch engine that replaces the reference's per-step
 *      `verlet` call (src/verlet.f90:65-1308) and its force loop, keeping
 *      state resident on the GPU between calls.
 */
#ifndef CARACAL_AMD_H
#define CARACAL_AMD_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
enum {
    CAM_OK = 0,
    CAM_ERR_ARG = 1,          /* invalid argument (null pointer, size out of range) */
    CAM_ERR_DEVICE = 2,       /* a HIP runtime call failed */
    CAM_ERR_NOMEM = 3,        /* device allocation failed */
    CAM_ERR_UNSUPPORTED = 4,  /* configuration the engine does not implement */
    CAM_ERR_STATE = 5         /* call out of order (e.g. step before set_state) */
};

/* Message of the last failure on this thread ("" if none). */
const char *cam_last_error(void);
/* Library version string. */
const char *cam_version(void);

/* ---- (1) analytic surfaces: drop-in for egrad_X(q,Natoms,Nbeads,V,dVdq,info) ----
 * q(3,natoms,nbeads) in bohr; V(nbeads) in hartree; dVdq(3,natoms,nbeads) in
 * hartree/bohr.  Host pointers; the call is synchronous (it copies over PCIe).
 * Replaces: egrad_h3 (src/egrad_h3.f:29-77), called per bead from gradient
 * (src/gradient.f90:181).  natoms must be 3. */
int cam_egrad_h3(const double *q, int natoms, int nbeads, double *v, double *dvdq);
/* Replaces: egrad_ch4h (src/egrad_ch4h.f:74-129), initialised by
 * initialize_ch4h (src/read_pes.f90:551).  natoms must be 6, atom order of the
 * reference surface: h4, C, h1, h2, h3, H_attacking. */
int cam_egrad_ch4h(const double *q, int natoms, int nbeads, double *v, double *dvdq);
/* Replaces: egrad_oh3 (src/egrad_oh3.f:33-190, OH + H2 of Schatz and
 * Elgersma), initialised by initialize_oh3 (src/read_pes.f90:491).  natoms
 * must be 4, atom order O, H1 (hydroxyl), H2, H3.  Every bead gets its own
 * gradient (the reference clears the whole array inside its bead loop and is
 * only called one bead at a time). */
int cam_egrad_oh3(const double *q, int natoms, int nbeads, double *v, double *dvdq);
/* Replaces: egrad_brh2 (src/egrad_brh2.f:33-79, the BrH2 DIM-3C surface of
 * Clary, Last and Baer), initialised by initialize_brh2 (src/read_pes.f90:452)
 * and called per bead from gradient (src/gradient.f90:182-183).  natoms must
 * be 3, atom order Ha, Br, Hc (the reference's R(1) = Ha-Br, R(2) = Ha-Hc,
 * R(3) = Br-Hc).  The 4 x 4 DIM eigenproblem is solved by Jacobi rotations
 * instead of the reference's EISPACK: energies agree to rounding, gradients
 * to the eigenvector's conditioning (within 1e-10 at the geometries checked). */
int cam_egrad_brh2(const double *q, int natoms, int nbeads, double *v, double *dvdq);
/* Replaces: egrad_o3 (src/egrad_o3.f:33-56, the O + O2 surface of Varga,
 * Paukku and Truhlar: permutationally invariant polynomials of
 * mixed exponential-Gaussian variables, a diatomic curve and D3(BJ)
 * dispersion), selected by ANA_O3 (src/read_pes.f90:456-473, no
 * initialisation call) and called per bead from gradient
 * (src/gradient.f90:184-185).  natoms must be 3 (three O atoms). */
int cam_egrad_o3(const double *q, int natoms, int nbeads, double *v, double *dvdq);

/* Device-resident variants: q, v, dvdq are device pointers, stream is a
 * hipStream_t (NULL = default stream); the call is asynchronous on that
 * stream.  precision selects CAM_PREC_FP64 or CAM_PREC_MIXED (FP32
 * arithmetic, FP64 inputs/outputs and FP64 accumulation). */
int cam_egrad_h3_dev(const double *q, int natoms, int nbeads, double *v, double *dvdq,
                     int precision, void *stream);
int cam_egrad_ch4h_dev(const double *q, int natoms, int nbeads, double *v, double *dvdq,
                       int precision, void *stream);
/* OH3, BrH2 and O3 have no binary32 form: precision must be CAM_PREC_FP64. */
int cam_egrad_oh3_dev(const double *q, int natoms, int nbeads, double *v, double *dvdq,
                      int precision, void *stream);
int cam_egrad_brh2_dev(const double *q, int natoms, int nbeads, double *v, double *dvdq,
                       int precision, void *stream);
int cam_egrad_o3_dev(const double *q, int natoms, int nbeads, double *v, double *dvdq,
                     int precision, void *stream);

/* Allocation-free form of cam_egrad_h3 / _ch4h / _oh3 / _brh2 / _o3 (the
 * surface by its CAM_SURFACE_* code; same arguments, host pointers,
 * synchronous).  No counterpart in the reference: it is for callers that
 * cannot allocate device memory per call (a tight per-step force loop, code
 * that forbids hipMalloc).  Its staging buffers live in one device array per
 * GPU, held until the process ends (4.9 MB); nbeads is limited to 16384
 * (CAM_ERR_ARG beyond).  Calls from several host threads on one GPU are
 * serialised. */
int cam_egrad_noalloc(int surface, const double *q, int natoms, int nbeads, double *v, double *dvdq);

/* ---- (2) trajectory-batch RPMD engine ---- */
enum { CAM_SURFACE_H3 = 1, CAM_SURFACE_CH4H = 2, CAM_SURFACE_OH3 = 3, CAM_SURFACE_BRH2 = 4, CAM_SURFACE_O3 = 5 };
/* Precision of the engine.  MIXED: forces in FP32 arithmetic with FP64
 * positions, momenta, accumulation and normal modes (CH4H: the whole
 * surface; H3: its asymmetric and compact correction terms, as
 * cam_egrad_h3_dev).  The OH3, BrH2 and O3 surfaces run in FP64 only
 * (CAM_PREC_MIXED is rejected). */
enum { CAM_PREC_FP64 = 0, CAM_PREC_MIXED = 1 };
/* Normal-mode transform.  REFERENCE reproduces the reference's rfft/irfft pair
 * (src/rfft.f90:36-62, src/irfft.f90:36-63): both directions are the
 * symmetric cosine matrix cos(2 pi k n / N)/sqrt(N), which projects the beads
 * onto their bead-reversal-symmetric part every step.  ORTHOGONAL is the
 * textbook orthogonal real-DFT normal-mode basis. */
enum { CAM_NM_REFERENCE = 0, CAM_NM_ORTHOGONAL = 1 };
/* What happens around the force evaluation in one step:
 *  NVE      -- dynamic.x without thermostat (verlet constrain = -1);
 *  ANDERSEN -- dynamic.x with the Andersen thermostat (thermostat = 1):
 *              all bead momenta redrawn every andersen_step steps;
 *  CHILD    -- recrossing child trajectory (verlet constrain = 2): the
 *              reaction coordinate of the centroid is evaluated each step and
 *              the recrossing-factor numerator is accumulated;
 *  PARENT   -- recrossing parent trajectory (verlet constrain = 1,
 *              src/recross_serial.f90:121-133, :284-299): the centroid is held
 *              on the dividing surface (constrain_q before the forces,
 *              constrain_p after the second half kick) and, if
 *              andersen_step > 0, the momenta are redrawn every andersen_step
 *              steps.  Runs on the split schedule; start it with
 *              cam_engine_begin_parent;
 *  UMBRELLA -- umbrella-sampling window trajectories (verlet with constrain
 *              = 0 in the sampling loops of src/calc_rate.f90:1556-1645):
 *              after the forces, the umbrella bias of the centroid
 *              (umbrella mode 0, src/umbrella.f90:136-173) is added to every
 *              bead's gradient, with each trajectory's own window
 *              (cam_engine_set_windows); the running sums of xi (mode 1)
 *              give each trajectory's average and variance
 *              (cam_engine_umbrella_stats).  Andersen as PARENT.  The fused
 *              schedule evaluates the bias inside the step kernel (one lane
 *              per trajectory, vector-ALU normal modes); the split schedule
 *              (or nm_kernel MFMA) in a kernel of its own after the force
 *              kernel.  Start it with cam_engine_begin_umbrella;
 *  NHC      -- dynamic.x with the Nose-Hoover chain thermostat (thermostat
 *              = 2, src/nhc.f90:35-169): four thermostats per trajectory,
 *              applied to all bead momenta before the first and after the
 *              second half kick (src/verlet.f90:185-206, :1086-1106), with
 *              the damping time nose_q; cam_engine_set_state resets them as
 *              mdinit does.  Fused schedule, vector-ALU normal modes. */
enum { CAM_ENS_NVE = 0, CAM_ENS_ANDERSEN = 1, CAM_ENS_CHILD = 2, CAM_ENS_PARENT = 3, CAM_ENS_UMBRELLA = 4,
       CAM_ENS_NHC = 5 };
/* Which unit computes the normal-mode transform inside the step kernel:
 * VALU (per-lane dot products over the beads) or MFMA (FP64 matrix-core tiles,
 * nbeads 16, 32 or 64).  FOLDED: the reference transform (rfft/irfft, the
 * real part of the DFT both ways) maps every ring polymer onto one that is
 * symmetric under the bead reflection j -> nbeads - j, so a step only has
 * nbeads/2 + 1 distinct beads; this kernel carries those, evaluates their
 * forces only, and writes each bead and its mirror (nm_mode REFERENCE,
 * nbeads >= 4, NVE or CHILD, fused schedule, no remove_com).  AUTO takes
 * FOLDED where it applies and the vector-ALU form otherwise, unless the
 * autotuner measures another faster.  All give the same trajectory to
 * rounding. */
enum { CAM_NMK_AUTO = 0, CAM_NMK_VALU = 1, CAM_NMK_MFMA = 2, CAM_NMK_FOLDED = 3 };
/* How a step is launched.  FUSED: the whole step, forces included, in one
 * kernel that advances steps_per_launch steps with the state in registers.
 * SPLIT: per step a normal-mode kernel and a force kernel (each at its own
 * register budget), the state in HBM in between.  AUTO: FUSED unless the
 * autotuner measures SPLIT faster.  Both give the same trajectory to
 * rounding. */
enum { CAM_SCHED_AUTO = 0, CAM_SCHED_FUSED = 1, CAM_SCHED_SPLIT = 2 };

#define CAM_MAX_ATOMS 16
#define CAM_MAX_BONDS 4
#define CAM_MAX_REAC 4

/* Bimolecular reaction coordinate (calc_xi BIMOLEC, src/calc_xi.f90:63-523;
 * keys of the reference's MECHA section, src/calc_rate_read.f90). 1-based atom
 * indices as in the reference key file. */
typedef struct {
    int32_t n_form, n_break;
    int32_t bond_form[CAM_MAX_BONDS][2];
    int32_t bond_break[CAM_MAX_BONDS][2];
    double form_ref[CAM_MAX_BONDS];      /* bond lengths at the transition state, bohr */
    double break_ref[CAM_MAX_BONDS];
    int32_t n_reac;                      /* number of reactant fragments (sum_reacs) */
    int32_t reac_natoms[CAM_MAX_REAC];
    int32_t reac_atoms[CAM_MAX_REAC][CAM_MAX_ATOMS];
    double r_inf;                        /* asymptotic reactant distance (dist_inf), bohr */
    double xi_ideal;                     /* dividing-surface value (recrossing uses mode 2) */
} cam_reaction;

typedef struct {
    int32_t surface;          /* CAM_SURFACE_* */
    int32_t natoms;           /* 3 (H3, BrH2, O3), 6 (CH4H) or 4 (OH3) */
    int32_t nbeads;           /* 1, 2, 4, 6, 8, 12, 16, 24, 32, 48 or 64 */
    int32_t ntraj;            /* trajectories in the batch */
    int32_t precision;        /* CAM_PREC_* */
    int32_t nm_mode;          /* CAM_NM_* */
    int32_t ensemble;         /* CAM_ENS_* */
    int32_t remove_com;       /* 1: remove centre-of-mass translation and rotation after every
                                 step as the reference's transrot does (src/transrot.f90) */
    int32_t andersen_step;    /* ANDERSEN, PARENT: resample every andersen_step steps (PARENT: 0 = never) */
    int32_t steps_per_launch; /* steps fused into one kernel launch; 0 = pick automatically */
    int32_t nm_kernel;        /* CAM_NMK_* */
    int32_t waves_per_eu;     /* register budget of the step kernel: 1 = up to 512 registers per lane,
                                 2 = 256 (two waves per SIMD; FP64 CH4H only); 0 = 2 for the FP64 CH4H
                                 folded kernel, 1 otherwise, unless tuned */
    int32_t schedule;         /* CAM_SCHED_* */
    int32_t reserved;         /* must be 0 */
    double dt;                /* time step, atomic time units */
    double mass[CAM_MAX_ATOMS];   /* electron masses */
    uint64_t seed;            /* counter-based RNG key for Andersen draws */
    cam_reaction reaction;    /* used by CAM_ENS_CHILD, CAM_ENS_PARENT and cam_engine_constrain */
    double nose_q;            /* NHC: thermostat damping time in atomic time units (the key file's
                                 NOSE_DAMP times the time step, src/dynamic.f90:578) */
} cam_config;

typedef struct cam_engine cam_engine;

int cam_engine_create(const cam_config *cfg, cam_engine **out);
int cam_engine_destroy(cam_engine *eng);

/* Upload state.  q, p: ntraj blocks of (3,natoms,nbeads), host memory;
 * beta: ntraj inverse temperatures 1/(k_B T) in 1/hartree (may be NULL only
 * for NVE).  Computes the forces of every bead (as mdinit does,
 * src/mdinit.f90:79-83). */
int cam_engine_set_state(cam_engine *eng, const double *q, const double *p, const double *beta);
int cam_engine_get_state(cam_engine *eng, double *q, double *p);
/* Gradient dV/dq of the current state, (3,natoms,nbeads) per trajectory, and
 * per-bead potential energies (nbeads per trajectory). */
int cam_engine_get_forces(cam_engine *eng, double *dvdq, double *vbead);

/* Redraw all bead momenta from the Maxwell-Boltzmann distribution at the
 * bead temperature (Andersen, src/andersen.f90:37-74).  'draw' selects the
 * counter of the stream (two calls with the same draw give the same
 * momenta).  If pair_antithetic != 0, trajectory 2m+1 receives minus the
 * momenta of trajectory 2m (the reference spawns children in +-p pairs,
 * src/recross_serial.f90:178-188). */
int cam_engine_draw_momenta(cam_engine *eng, uint64_t draw, int pair_antithetic);

/* Shift every trajectory's beads so its centroid lies on the dividing
 * surface (constrain_q, src/constrain_q.f90:35-113), then project the
 * momenta (constrain_p, src/constrain_p.f90:36-76) if project_p != 0.
 * status (ntraj ints, may be NULL) receives 0 on success, 1 if the Newton
 * iteration did not converge in 200 steps. */
int cam_engine_constrain(cam_engine *eng, int project_p, int *status);

/* Start a recrossing child batch: evaluates xi and its gradient at the
 * centroid and the flux factor vs/fs of every trajectory
 * (src/recross_serial.f90:189-233) and adds vs/fs to the denominator where
 * vs > 0.  The following cam_engine_run accumulates the numerator num(t)
 * step by step; like the reference's num_total/denom_total, both keep
 * accumulating over successive child batches of the same engine. */
int cam_engine_begin_children(cam_engine *eng);

/* Start (or restart) recrossing parents, as mdinit with bias_mode 2 does
 * (src/mdinit.f90:79-98, called at src/recross_serial.f90:118 and :270):
 * fresh forces of every bead, plus the umbrella bias gradient of the
 * centroid (src/umbrella.f90:136-173: k_force (xi - xi_ideal) dxi with xi of
 * mode 1, and the Hessian metric term) on every bead; the gradient of xi
 * (mode 1) becomes the constraint direction of the first step.  Needed after
 * every cam_engine_set_state of a PARENT engine before it runs.  mdinit's
 * momentum draw is cam_engine_draw_momenta.
 * It also arms the per-step trajectory check of the recrossing driver
 * (rpmd_check, src/rpmd_check.f90:74-98, as src/recross_serial.f90:125-135
 * and :290-299 call it): from step check_after + 1 after this call on (never
 * if check_after < 0), a trajectory whose positions hold a NaN or a value
 * above HUGE, or whose energy (sum over the beads, plus the reference's 1e5
 * penalty in a step whose constraint failed) is NaN, above HUGE or above
 * epot_max, is flagged (cam_engine_observables).  The reference's limit is
 * (ts_energy + energy_tol) * nbeads; INFINITY keeps only the NaN and HUGE
 * checks.  Clears the flags. */
int cam_engine_begin_parent(cam_engine *eng, double k_force, double epot_max, int check_after);

/* UMBRELLA: the window of every trajectory, xi_ideal[t] (the xi_wins value,
 * mode 1) and k_force[t] (hartree), host arrays of ntraj.  Until called,
 * every window sits at the reaction's xi_ideal with k_force 0.  Needs a new
 * cam_engine_begin_umbrella. */
int cam_engine_set_windows(cam_engine *eng, const double *xi_ideal, const double *k_force);

/* UMBRELLA: start a segment of window steps.  bias != 0 is mdinit's start
 * (src/mdinit.f90:79-98): fresh forces plus the umbrella bias of each
 * window, as before the equilibration (src/calc_rate.f90:1549); bias = 0
 * gives the unbiased forces the reference starts its sampling from
 * (src/calc_rate.f90:1610-1618).  Clears the running sums of xi, the flags
 * and the step count of the Andersen schedule.  Arms the check of the
 * umbrella loops (rpmd_check with traj_type 2 and 3, src/rpmd_check.f90:74-112)
 * as cam_engine_begin_parent does, with xi_tol added: |xi - xi_ideal| > xi_tol
 * fails (INFINITY disables that part). */
int cam_engine_begin_umbrella(cam_engine *eng, int bias, double epot_max, double xi_tol, int check_after);

/* UMBRELLA: average and variance of xi (mode 1) of every trajectory over the
 * steps run since cam_engine_begin_umbrella, as src/calc_rate.f90:1639-1657
 * forms them (sum/n, sum of squares/n - average^2).  Host arrays of ntraj. */
int cam_engine_umbrella_stats(cam_engine *eng, double *average, double *variance);

/* UMBRELLA: histograms of the same samples, for WHAM (src/wham.f90, which
 * calc_rate feeds with per-window histograms of xi, src/calc_rate.f90:1793-1840;
 * the reference's writer of those histograms is commented out,
 * src/calc_rate.f90:1712-1724, so they are collected here on the device).
 * Every step's xi counts into bin j = nint((xi - xi_min) / bin_size),
 * 1 <= j <= nbins, the bin src/wham.f90:104-106 gives the coordinate
 * xi_min + j bin_size; samples outside are not counted.  nbins = 0 turns the
 * histograms off.  cam_engine_begin_umbrella clears them.  Captured graphs
 * are dropped.  No counterpart in the reference's engine. */
int cam_engine_set_umbrella_histogram(cam_engine *eng, int nbins, double xi_min, double bin_size);
/* counts[traj * nbins + j - 1]: samples of trajectory traj in bin j since
 * cam_engine_begin_umbrella (host array of ntraj * nbins). */
int cam_engine_umbrella_histogram(cam_engine *eng, double *counts);

/* Advance every trajectory by nsteps velocity-Verlet RPMD steps.  Runs on the
 * engine's stream; all steps are captured in one HIP graph on first use and
 * replayed afterwards.  Asynchronous: call cam_engine_sync to wait.  The
 * Andersen schedule of a PARENT counts steps from cam_engine_begin_parent
 * (first step 1), as the reference's verlet loop restarts after mdinit; its
 * draws are keyed by the steps since creation (first step 1). */
int cam_engine_run(cam_engine *eng, int nsteps);
/* Capture and instantiate the graph of an nsteps run without running it, so
 * a later cam_engine_run(nsteps) only replays it. */
int cam_engine_prepare(cam_engine *eng, int nsteps);
int cam_engine_sync(cam_engine *eng);

/* Recrossing factor kappa(t) = num(t)/den.  num(t) counts child step t since
 * the last cam_engine_begin_children (successive cam_engine_run calls
 * continue the same children) and, like den, accumulates over successive
 * child batches.  Host array of nsteps doubles for t = 1..nsteps; den is
 * returned separately.  Sums run over the batch in a fixed order, so the
 * result does not depend on scheduling. */
int cam_engine_kappa(cam_engine *eng, double *num, double *den, int nsteps);

/* Per-trajectory observables of the current state (host arrays of ntraj):
 * epot = sum over beads of V (as verlet returns it), ekin = sum p^2/2m over
 * beads, espring = ring-polymer spring energy sum 1/2 m w_n^2 (q_k - q_k+1)^2
 * with w_n = nbeads/beta, xi = reaction coordinate of the centroid (mode 2;
 * CHILD, PARENT and UMBRELLA, else 0), flags = bit 0: non-finite position
 * now; since cam_engine_begin_parent / _begin_umbrella: bit 1 (PARENT) a
 * constraint did not converge, bit 2 the trajectory check failed. */
int cam_engine_observables(cam_engine *eng, double *epot, double *ekin, double *espring,
                           double *xi, int *flags);

/* Device pointers of the resident state (q, p, dvdq) in the engine's internal
 * layout [traj][3*natoms][nbeads] and the hipStream_t it runs on. */
int cam_engine_device_state(cam_engine *eng, double **q, double **p, double **dvdq, void **stream);

/* Schedule autotuner.  Times candidate schedules of the step kernel on a
 * copy of the current state (the engine's trajectories, kappa sums and RNG
 * counters are not touched) for trial_steps steps each, and keeps the fastest:
 * the normal-mode unit (VALU or MFMA; only when nm_kernel is AUTO), fused or
 * split launches (schedule AUTO, remove_com off), and for the fused schedule
 * the register budget (waves_per_eu 1 or 2; FP64 CH4H with waves_per_eu 0)
 * and the steps fused per launch (10, 25, 50, 100; only when
 * steps_per_launch is 0 and remove_com is off).  Captured graphs are dropped.  No counterpart in
 * the reference, which has one schedule. */
#define CAM_MAX_TRIALS 32
typedef struct {
    int32_t n;                              /* candidates timed */
    int32_t chosen;                         /* index of the schedule kept */
    int32_t nm_kernel[CAM_MAX_TRIALS];      /* CAM_NMK_VALU or CAM_NMK_MFMA */
    int32_t steps_per_launch[CAM_MAX_TRIALS];
    int32_t waves_per_eu[CAM_MAX_TRIALS];
    int32_t split[CAM_MAX_TRIALS];          /* 1: the split schedule */
    double us_per_step[CAM_MAX_TRIALS];     /* microseconds per step of the whole batch */
} cam_tuning;
int cam_engine_autotune(cam_engine *eng, int trial_steps, cam_tuning *out);

/* Timing of the last cam_engine_run (after cam_engine_sync): the GPU time of
 * the whole run (HIP events on the engine's stream around its graph
 * replays, so the small per-launch kernels are included) divided by the
 * number of step launches it enqueued, the steps one launch advances
 * (min(steps_per_launch, nsteps) fused, 1 split; a run whose nsteps is not a
 * multiple ends with a shorter launch) and the number of launches. */
int cam_engine_last_timing(cam_engine *eng, double *ms_per_launch, int *steps_per_launch, int *launches);

/* Ragged batches (trajectories of different surfaces and bead counts): one
 * engine per bucket of equal (surface, nbeads).  cam_batch_run replays one
 * captured graph per run length: forked from the batch stream, every
 * bucket's steps as a branch on its own stream (longest bucket first by a
 * work estimate, trajectories x beads x surface flops), joined back; the
 * buckets run concurrently on the GPU.  The graph is captured again when a
 * bucket's engine changed schedule or buffers.  Engines are reached with
 * cam_batch_engine to set state and read results; the batch owns them; a
 * bucket's engine must not be run on its own while a batch run is in
 * flight.  Batches hold NVE, Andersen, Nose-Hoover and child engines.  An
 * engine run on its own issues folded launches of at least 512 blocks as two
 * trajectory slices on two streams (the same bits); a bucket does not, and
 * cam_batch_autotune times each bucket's schedules that way.  No
 * counterpart in the reference, which runs one trajectory per MPI rank
 * (src/recross.f90:334-415). */
typedef struct cam_batch cam_batch;
int cam_batch_create(const cam_config *cfgs, int nbuckets, cam_batch **out);
int cam_batch_destroy(cam_batch *batch);
int cam_batch_engine(cam_batch *batch, int bucket, cam_engine **eng);
int cam_batch_autotune(cam_batch *batch, int trial_steps);
int cam_batch_run(cam_batch *batch, int nsteps);
/* Waits for the last run; ms = its GPU time from the first launch to the
 * last bucket's end (HIP events). */
int cam_batch_sync(cam_batch *batch, double *ms);

/* ---- (3) QMDFF-form periodic box: energy kernel, hand-written adjoint and
 * a ring-polymer engine ----
 * Terms follow the reference's QMDFF force field: bond stretch and bend of
 * ff_eg (src/ff_eg.f90:87-299), BJ-damped dispersion, Pauli repulsion and the
 * Zahn-damped Coulomb of ff_nonb (src/ff_nonb.f90:198-394) between atoms of
 * different molecules, minimum image as box_image (src/box_image.f90:36-64).
 * Forms: Morse De(1-exp(-a(r-r0)))^2 bonds, harmonic k(theta-theta0)^2
 * angles.  Two atom types (0, 1).  Parameter vector theta (10 doubles):
 * De, a, r0, k, theta0, q_0, q_1, C6_00, C6_01, C6_11. */
typedef struct {
    int32_t natoms, nbonds, nangles;
    const int32_t *type;      /* natoms, 0 or 1 */
    const int32_t *mol;       /* natoms, molecule index (nonbonded terms skip equal indices) */
    const int32_t *bonds;     /* 2*nbonds, 0-based */
    const int32_t *angles;    /* 3*nangles (i, vertex, k), 0-based */
    double box;               /* cubic box edge, bohr */
    double rcut;              /* nonbonded cutoff, bohr (<= box/2) */
    double zahn_a, zahn_par;  /* Zahn damping (src/set_periodic.f90:93-97) */
    double sr42[2][2], r0bj[2][2];   /* BJ dispersion: C8/C6 factor and damping radius */
    double zrep[2][2], arep[2][2];   /* repulsion prefactor and exponent */
} cam_ff_desc;

typedef struct cam_ff cam_ff;
int cam_ff_create(const cam_ff_desc *desc, cam_ff **out);
/* The reference's flexible water model (src/egrad_water.f90:36-333) as a
 * force field for the calls below and for cam_box_create: nwater molecules,
 * atoms O, H, H per molecule, parameters of water_init
 * (src/water_init.f90:83-106).  periodic: cubic box of side box (bohr),
 * minimum image, cutoff rcut (bohr, <= box/2) on the intermolecular terms;
 * zahn: Zahn-damped Coulomb with the constants of set_periodic
 * (src/set_periodic.f90:93-97).  The intermolecular gradient is the
 * reference's (e0/r^2 per pair) and LJ acts when the lower-index atom of a
 * pair is an oxygen, as in the reference.  theta is ignored (may be NULL);
 * parameter derivatives are not available. */
int cam_water_create(int nwater, int periodic, double box, int zahn, double rcut, cam_ff **out);
int cam_ff_destroy(cam_ff *ff);
/* How the QMDFF-form nonbonded pairs are found.  The atoms are kept in a
 * spatial order (cells of their bead centroids) that is rebuilt on the
 * device when some bead has moved more than half a 1.2 bohr skin since the
 * last build, or the bead count changed.  LIST (default): per atom a row of
 * every atom of another molecule whose centroid was within rc + skin + d_i +
 * d_j (d = the largest bead distance from its atom's centroid) at the
 * build, so that no pair inside the cutoff in any bead is missed.  ALL:
 * every pair tested, as the reference's loop (src/ff_nonb.f90:198-332).  Both
 * meet the pairs inside the cutoff in the same order and give the same
 * results bit for bit.  The mode applies to work enqueued afterwards (a box
 * engine's captured graphs keep the mode they were captured with).  Calls on
 * one force field from several streams at once race on its pair structure.
 * No counterpart in the reference, which has no neighbour list. */
enum { CAM_FF_PAIRS_LIST = 0, CAM_FF_PAIRS_ALL = 1 };
int cam_ff_set_pair_mode(cam_ff *ff, int mode);
/* Energy of nbeads replicas: x device [nbeads][natoms][3] bohr, e device
 * [nbeads] hartree.  theta is a host array of 10.  Asynchronous on stream. */
int cam_ff_energy(cam_ff *ff, const double *theta, const double *x, int nbeads, double *e, int precision,
                  void *stream);
/* Adjoint of the energy kernel: dE/dx (device, same layout as x); optionally
 * e [nbeads], dE/dtheta [nbeads][10] and, for a direction v (device, layout
 * of x), d(v . dE/dx)/dtheta [nbeads][10] -- the parameter gradient of a
 * force-matching loss.  No counterpart in the reference, which fits QMDFF
 * parameters without derivatives (src/lm_qmdff.f90:36-84). */
int cam_ff_forces(cam_ff *ff, const double *theta, const double *x, int nbeads, double *e, double *dedx,
                  double *dedtheta, const double *v, double *dvdtheta, int precision, void *stream);

/* The reference's own QMDFF force field (one QMDFF, src/gradient.f90:341-348:
 * ff_eg + ff_nonb + ff_hb) as a cam_ff for cam_ff_energy / cam_ff_forces and
 * cam_box_create.  The descriptor holds what the reference's prepare
 * (src/prepare.f90:30-196: rdsolvff, setnonb, copyc6, ncoord_qmdff, getc6)
 * and set_periodic (src/set_periodic.f90) leave in the qmdff, evb_mod and
 * pbc_mod modules, in their layouts: 1-based atom indices, per-term
 * parameter columns, 94 x 94 element-pair tables column-major (element
 * (z_a, z_b) at [(z_b - 1) * 94 + z_a - 1]), lengths in bohr.  Terms: bond
 * stretch k(1 + (r0/r)^a - 2 (r0/r)^(a/2)), bends k(cos t - cos t0)^2 (or
 * k(t - t0)^2 for t0 ~ pi) damped by abdamp, torsions and out-of-plane terms
 * (src/ff_eg.f90:87-619), the nci pair list with eps2/eps1 scaling and the
 * pairs between molecules (dispersion, repulsion, Coulomb plain, switched or
 * Zahn; src/ff_nonb.f90:88-509), listed and intermolecular hydrogen bonds
 * (src/ff_hb.f90:60-276, src/eabhag.f90).  Halogen bonds (eabxag) and Ewald
 * sums are not implemented (CAM_ERR_UNSUPPORTED).  Energies exclude the
 * QMDFF shift (E_zero1).  theta is ignored; no parameter derivatives. */
typedef struct {
    int32_t natoms, nbond, nangl, ntors, nhb, nnci;
    const int32_t *at;        /* natoms atomic numbers 1..94 */
    const double *molnum;     /* natoms molecule numbers (all 0: one molecule) */
    const double *q;          /* natoms charges */
    const int32_t *bond;      /* 2 * nbond: bond(1:2, m) */
    const double *vbond;      /* 3 * nbond: r0, k, exponent */
    const int32_t *angl;      /* 3 * nangl: vertex, then the outer atoms */
    const double *vangl;      /* 2 * nangl: theta0, k */
    const int32_t *tors;      /* 6 * ntors: i, j, k, l, number of terms, kind (2: out of plane) */
    const double *vtors;      /* 14 * ntors */
    const int32_t *hb;        /* 3 * nhb: A, B, H */
    const double *vhb;        /* 2 * nhb */
    const int32_t *nci;       /* 3 * nnci: i, j, kind 1..6 */
    const double *c6xy;       /* natoms * natoms, symmetric */
    const double *r094, *sr42, *zab, *r0ab;   /* 94 x 94 each */
    double eps1[6], eps2[6];
    const double *rad;        /* 94 covalent radii, bohr */
    const double *scalehb;    /* 94 hydrogen-bond scaling factors */
    int32_t periodic, zahn;
    double box[3];            /* box lengths, bohr */
    double coul_cut, vdw_cut, cut_low, zahn_a, zahn_par;
} cam_qmdff_desc;
int cam_qmdff_create(const cam_qmdff_desc *desc, cam_ff **out);

/* Ring-polymer dynamics of one box (replaces verlet's per-bead force loop
 * over gradient -> ff_eg/ff_nonb, src/verlet.f90:772-777, for a QMDFF box).
 * andersen_step 0 = NVE.  State arrays are host (3,natoms,nbeads). */
typedef struct cam_box cam_box;
int cam_box_create(cam_ff *ff, int nbeads, double dt, const double *mass, double beta, int nm_mode, int precision,
                   int andersen_step, uint64_t seed, cam_box **out);
int cam_box_destroy(cam_box *box);
int cam_box_set_params(cam_box *box, const double *theta);
int cam_box_set_state(cam_box *box, const double *q, const double *p);
int cam_box_get_state(cam_box *box, double *q, double *p);
int cam_box_prepare(cam_box *box, int nsteps);
int cam_box_run(cam_box *box, int nsteps);
/* Nose-Hoover chain thermostat for the box (thermostat 2, src/nhc.f90:35-169,
 * applied before the first and after the second half kick of every step as
 * src/verlet.f90:185-206 and :1086-1106 do): nose_q is the damping time in
 * atomic time units (NOSE_DAMP x dt, src/dynamic.f90:578), 0 turns it off.
 * Needs andersen_step 0.  The chain starts at rest, and again after every
 * cam_box_set_state (mdinit). */
int cam_box_set_nhc(cam_box *box, double nose_q);
/* Centre-of-mass translation and rotation removal after every step, as
 * dynamic.x's verlet ends each step with transrot (src/verlet.f90:1299-1305,
 * src/transrot.f90:36-237, its total mass nbeads^2 * M included); on != 0
 * turns it on.  Off by default. */
int cam_box_set_transrot(cam_box *box, int on);
/* Waits for the last run; ms (may be NULL) receives its duration measured
 * with HIP events on the box's stream. */
int cam_box_sync(cam_box *box, double *ms);
/* Per-bead potential energies (nbeads), total kinetic and spring energy. */
int cam_box_energies(cam_box *box, double *epot_beads, double *ekin, double *espring);
int cam_box_device_state(cam_box *box, double **q, double **p, double **dvdq, void **stream);

#ifdef __cplusplus
}
#endif
#endif
