!
!   caracal_amd_mod: iso_c_binding interface of the MI355X RPMD engine
!   (include/caracal_amd.h) for a Fortran host such as Caracal's dynamic.x
!   and calc_rate.x.  Every function returns the integer status of the C
!   entry point (0 = success); cam_error_message returns the text of the
!   last failure.
!
!   Arrays keep the reference layout: q(3,natoms,nbeads) per trajectory, a
!   batch is q(3,natoms,nbeads,ntraj).  Units: bohr, hartree, electron
!   masses, atomic time units.
!
module caracal_amd_mod
   use, intrinsic :: iso_c_binding
   implicit none
   private

   integer(c_int), parameter, public :: CAM_OK = 0
   integer(c_int32_t), parameter, public :: CAM_SURFACE_H3 = 1, CAM_SURFACE_CH4H = 2, CAM_SURFACE_OH3 = 3, &
                                            CAM_SURFACE_BRH2 = 4, CAM_SURFACE_O3 = 5
   integer(c_int32_t), parameter, public :: CAM_PREC_FP64 = 0, CAM_PREC_MIXED = 1
   integer(c_int32_t), parameter, public :: CAM_NM_REFERENCE = 0, CAM_NM_ORTHOGONAL = 1
   integer(c_int32_t), parameter, public :: CAM_ENS_NVE = 0, CAM_ENS_ANDERSEN = 1, CAM_ENS_CHILD = 2, &
                                             CAM_ENS_PARENT = 3, CAM_ENS_UMBRELLA = 4, CAM_ENS_NHC = 5
   integer(c_int32_t), parameter, public :: CAM_NMK_AUTO = 0, CAM_NMK_VALU = 1, CAM_NMK_MFMA = 2, &
                                            CAM_NMK_FOLDED = 3
   integer(c_int32_t), parameter, public :: CAM_SCHED_AUTO = 0, CAM_SCHED_FUSED = 1, CAM_SCHED_SPLIT = 2
   integer(c_int32_t), parameter, public :: CAM_FF_PAIRS_LIST = 0, CAM_FF_PAIRS_ALL = 1
   integer, parameter, public :: CAM_MAX_ATOMS = 16, CAM_MAX_BONDS = 4, CAM_MAX_REAC = 4

   ! cam_reaction: the BIMOLEC reaction coordinate (MECHA section keys)
   type, bind(C), public :: cam_reaction
      integer(c_int32_t) :: n_form = 0, n_break = 0
      integer(c_int32_t) :: bond_form(2, CAM_MAX_BONDS) = 0
      integer(c_int32_t) :: bond_break(2, CAM_MAX_BONDS) = 0
      real(c_double) :: form_ref(CAM_MAX_BONDS) = 0.d0
      real(c_double) :: break_ref(CAM_MAX_BONDS) = 0.d0
      integer(c_int32_t) :: n_reac = 0
      integer(c_int32_t) :: reac_natoms(CAM_MAX_REAC) = 0
      integer(c_int32_t) :: reac_atoms(CAM_MAX_ATOMS, CAM_MAX_REAC) = 0
      real(c_double) :: r_inf = 0.d0, xi_ideal = 0.d0
   end type cam_reaction

   type, bind(C), public :: cam_config
      integer(c_int32_t) :: surface = CAM_SURFACE_H3, natoms = 3, nbeads = 1, ntraj = 1
      integer(c_int32_t) :: precision = CAM_PREC_FP64, nm_mode = CAM_NM_REFERENCE
      integer(c_int32_t) :: ensemble = CAM_ENS_NVE, remove_com = 0, andersen_step = 0
      integer(c_int32_t) :: steps_per_launch = 0, nm_kernel = 0, waves_per_eu = 0, schedule = 0, reserved = 0
      real(c_double) :: dt = 0.d0
      real(c_double) :: mass(CAM_MAX_ATOMS) = 0.d0
      integer(c_int64_t) :: seed = 0
      type(cam_reaction) :: reaction
      real(c_double) :: nose_q = 0.0_c_double
   end type cam_config

   ! cam_qmdff_desc: what prepare and set_periodic leave in the qmdff,
   ! evb_mod and pbc_mod modules (include/caracal_amd.h); the pointers are
   ! c_loc of those arrays, which must stay allocated while cam_qmdff_create
   ! runs (it copies them)
   type, bind(C), public :: cam_qmdff_desc
      integer(c_int32_t) :: natoms = 0, nbond = 0, nangl = 0, ntors = 0, nhb = 0, nnci = 0
      type(c_ptr) :: at = c_null_ptr, molnum = c_null_ptr, q = c_null_ptr
      type(c_ptr) :: bond = c_null_ptr, vbond = c_null_ptr, angl = c_null_ptr, vangl = c_null_ptr
      type(c_ptr) :: tors = c_null_ptr, vtors = c_null_ptr, hb = c_null_ptr, vhb = c_null_ptr
      type(c_ptr) :: nci = c_null_ptr, c6xy = c_null_ptr
      type(c_ptr) :: r094 = c_null_ptr, sr42 = c_null_ptr, zab = c_null_ptr, r0ab = c_null_ptr
      real(c_double) :: eps1(6) = 0.d0, eps2(6) = 0.d0
      type(c_ptr) :: rad = c_null_ptr, scalehb = c_null_ptr
      integer(c_int32_t) :: periodic = 0, zahn = 0
      real(c_double) :: box(3) = 0.d0
      real(c_double) :: coul_cut = 0.d0, vdw_cut = 0.d0, cut_low = 0.d0, zahn_a = 0.d0, zahn_par = 0.d0
   end type cam_qmdff_desc

   integer, parameter, public :: CAM_MAX_TRIALS = 32
   type, bind(C), public :: cam_tuning
      integer(c_int32_t) :: n = 0, chosen = 0
      integer(c_int32_t) :: nm_kernel(CAM_MAX_TRIALS) = 0, steps_per_launch(CAM_MAX_TRIALS) = 0
      integer(c_int32_t) :: waves_per_eu(CAM_MAX_TRIALS) = 0, split(CAM_MAX_TRIALS) = 0
      real(c_double) :: us_per_step(CAM_MAX_TRIALS) = 0.d0
   end type cam_tuning

   interface
      function cam_last_error() bind(C, name="cam_last_error") result(msg)
         import :: c_ptr
         type(c_ptr) :: msg
      end function
      ! drop-in for egrad_h3(q,Natoms,Nbeads,V,dVdq,info) (src/egrad_h3.f:29)
      function cam_egrad_h3(q, natoms, nbeads, v, dvdq) bind(C, name="cam_egrad_h3") result(rc)
         import :: c_int, c_double
         real(c_double), intent(in) :: q(*)
         integer(c_int), value :: natoms, nbeads
         real(c_double), intent(out) :: v(*), dvdq(*)
         integer(c_int) :: rc
      end function
      ! drop-in for egrad_ch4h(q,Natoms,Nbeads,V,dVdq,info) (src/egrad_ch4h.f:74)
      function cam_egrad_ch4h(q, natoms, nbeads, v, dvdq) bind(C, name="cam_egrad_ch4h") result(rc)
         import :: c_int, c_double
         real(c_double), intent(in) :: q(*)
         integer(c_int), value :: natoms, nbeads
         real(c_double), intent(out) :: v(*), dvdq(*)
         integer(c_int) :: rc
      end function
      ! drop-in for egrad_oh3(q,Natoms,Nbeads,V,dVdq,info) (src/egrad_oh3.f:33)
      function cam_egrad_oh3(q, natoms, nbeads, v, dvdq) bind(C, name="cam_egrad_oh3") result(rc)
         import :: c_int, c_double
         real(c_double), intent(in) :: q(*)
         integer(c_int), value :: natoms, nbeads
         real(c_double), intent(out) :: v(*), dvdq(*)
         integer(c_int) :: rc
      end function
      ! drop-in for egrad_brh2(q,Natoms,Nbeads,V,dVdq,info) (src/egrad_brh2.f:33)
      function cam_egrad_brh2(q, natoms, nbeads, v, dvdq) bind(C, name="cam_egrad_brh2") result(rc)
         import :: c_int, c_double
         real(c_double), intent(in) :: q(*)
         integer(c_int), value :: natoms, nbeads
         real(c_double), intent(out) :: v(*), dvdq(*)
         integer(c_int) :: rc
      end function
      ! drop-in for egrad_o3(q,Natoms,Nbeads,V,dVdq,info) (src/egrad_o3.f:33)
      function cam_egrad_o3(q, natoms, nbeads, v, dvdq) bind(C, name="cam_egrad_o3") result(rc)
         import :: c_int, c_double
         real(c_double), intent(in) :: q(*)
         integer(c_int), value :: natoms, nbeads
         real(c_double), intent(out) :: v(*), dvdq(*)
         integer(c_int) :: rc
      end function
      function cam_engine_create(cfg, eng) bind(C, name="cam_engine_create") result(rc)
         import :: c_int, c_ptr, cam_config
         type(cam_config), intent(in) :: cfg
         type(c_ptr), intent(out) :: eng
         integer(c_int) :: rc
      end function
      function cam_engine_destroy(eng) bind(C, name="cam_engine_destroy") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: eng
         integer(c_int) :: rc
      end function
      function cam_engine_set_state(eng, q, p, beta) bind(C, name="cam_engine_set_state") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(in) :: q(*), p(*), beta(*)
         integer(c_int) :: rc
      end function
      function cam_engine_get_state(eng, q, p) bind(C, name="cam_engine_get_state") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(out) :: q(*), p(*)
         integer(c_int) :: rc
      end function
      function cam_engine_get_forces(eng, dvdq, vbead) bind(C, name="cam_engine_get_forces") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(out) :: dvdq(*), vbead(*)
         integer(c_int) :: rc
      end function
      function cam_engine_draw_momenta(eng, draw, pair) bind(C, name="cam_engine_draw_momenta") result(rc)
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: eng
         integer(c_int64_t), value :: draw
         integer(c_int), value :: pair
         integer(c_int) :: rc
      end function
      function cam_engine_constrain(eng, project_p, status) bind(C, name="cam_engine_constrain") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: eng
         integer(c_int), value :: project_p
         integer(c_int), intent(out) :: status(*)
         integer(c_int) :: rc
      end function
      function cam_engine_begin_children(eng) bind(C, name="cam_engine_begin_children") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: eng
         integer(c_int) :: rc
      end function
      function cam_engine_begin_parent(eng, k_force, epot_max, check_after) &
            bind(C, name="cam_engine_begin_parent") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), value :: k_force, epot_max
         integer(c_int), value :: check_after
         integer(c_int) :: rc
      end function
      function cam_engine_set_windows(eng, xi_ideal, k_force) bind(C, name="cam_engine_set_windows") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(in) :: xi_ideal(*), k_force(*)
         integer(c_int) :: rc
      end function
      function cam_engine_begin_umbrella(eng, bias, epot_max, xi_tol, check_after) &
            bind(C, name="cam_engine_begin_umbrella") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         integer(c_int), value :: bias, check_after
         real(c_double), value :: epot_max, xi_tol
         integer(c_int) :: rc
      end function
      function cam_engine_umbrella_stats(eng, average, variance) bind(C, name="cam_engine_umbrella_stats") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(out) :: average(*), variance(*)
         integer(c_int) :: rc
      end function
      ! per-window histograms of xi for wham (counts(nbins, ntraj))
      function cam_engine_set_umbrella_histogram(eng, nbins, xi_min, bin_size) &
            bind(C, name="cam_engine_set_umbrella_histogram") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         integer(c_int), value :: nbins
         real(c_double), value :: xi_min, bin_size
         integer(c_int) :: rc
      end function
      function cam_engine_umbrella_histogram(eng, counts) bind(C, name="cam_engine_umbrella_histogram") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(out) :: counts(*)
         integer(c_int) :: rc
      end function
      function cam_engine_run(eng, nsteps) bind(C, name="cam_engine_run") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: eng
         integer(c_int), value :: nsteps
         integer(c_int) :: rc
      end function
      function cam_engine_sync(eng) bind(C, name="cam_engine_sync") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: eng
         integer(c_int) :: rc
      end function
      function cam_engine_kappa(eng, num, den, nsteps) bind(C, name="cam_engine_kappa") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(out) :: num(*), den
         integer(c_int), value :: nsteps
         integer(c_int) :: rc
      end function
      function cam_engine_observables(eng, epot, ekin, espring, xi, flags) &
            bind(C, name="cam_engine_observables") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(out) :: epot(*), ekin(*), espring(*), xi(*)
         integer(c_int), intent(out) :: flags(*)
         integer(c_int) :: rc
      end function
      function cam_engine_prepare(eng, nsteps) bind(C, name="cam_engine_prepare") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: eng
         integer(c_int), value :: nsteps
         integer(c_int) :: rc
      end function
      function cam_engine_autotune(eng, trial_steps, tuning) bind(C, name="cam_engine_autotune") result(rc)
         import :: c_int, c_ptr, cam_tuning
         type(c_ptr), value :: eng
         integer(c_int), value :: trial_steps
         type(cam_tuning), intent(out) :: tuning
         integer(c_int) :: rc
      end function
      function cam_engine_last_timing(eng, ms_per_launch, steps_per_launch, launches) &
            bind(C, name="cam_engine_last_timing") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(out) :: ms_per_launch
         integer(c_int), intent(out) :: steps_per_launch, launches
         integer(c_int) :: rc
      end function
      ! ragged batches: one engine per (surface, nbeads) bucket, run concurrently
      function cam_batch_create(cfgs, nbuckets, batch) bind(C, name="cam_batch_create") result(rc)
         import :: c_int, c_ptr, cam_config
         type(cam_config), intent(in) :: cfgs(*)
         integer(c_int), value :: nbuckets
         type(c_ptr), intent(out) :: batch
         integer(c_int) :: rc
      end function
      function cam_batch_destroy(batch) bind(C, name="cam_batch_destroy") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: batch
         integer(c_int) :: rc
      end function
      function cam_batch_engine(batch, bucket, eng) bind(C, name="cam_batch_engine") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: batch
         integer(c_int), value :: bucket
         type(c_ptr), intent(out) :: eng
         integer(c_int) :: rc
      end function
      function cam_batch_autotune(batch, trial_steps) bind(C, name="cam_batch_autotune") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: batch
         integer(c_int), value :: trial_steps
         integer(c_int) :: rc
      end function
      function cam_batch_run(batch, nsteps) bind(C, name="cam_batch_run") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: batch
         integer(c_int), value :: nsteps
         integer(c_int) :: rc
      end function
      function cam_batch_sync(batch, ms) bind(C, name="cam_batch_sync") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: batch
         real(c_double), intent(out) :: ms
         integer(c_int) :: rc
      end function
      ! the reference's water model as a force field (src/egrad_water.f90)
      function cam_water_create(nwater, periodic, box, zahn, rcut, ff) bind(C, name="cam_water_create") result(rc)
         import :: c_int, c_ptr, c_double
         integer(c_int), value :: nwater, periodic, zahn
         real(c_double), value :: box, rcut
         type(c_ptr), intent(out) :: ff
         integer(c_int) :: rc
      end function
      function cam_ff_destroy(ff) bind(C, name="cam_ff_destroy") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: ff
         integer(c_int) :: rc
      end function
      function cam_ff_set_pair_mode(ff, mode) bind(C, name="cam_ff_set_pair_mode") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: ff
         integer(c_int), value :: mode
         integer(c_int) :: rc
      end function
      ! the reference's QMDFF (ff_eg + ff_nonb + ff_hb) as a force field
      function cam_qmdff_create(desc, ff) bind(C, name="cam_qmdff_create") result(rc)
         import :: c_int, c_ptr, cam_qmdff_desc
         type(cam_qmdff_desc), intent(in) :: desc
         type(c_ptr), intent(out) :: ff
         integer(c_int) :: rc
      end function
      ! ring-polymer dynamics of one box; q, p as q(3,natoms,nbeads)
      function cam_box_create(ff, nbeads, dt, mass, beta, nm_mode, precision, andersen_step, seed, box) &
            bind(C, name="cam_box_create") result(rc)
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: ff
         integer(c_int), value :: nbeads, nm_mode, precision, andersen_step
         real(c_double), value :: dt, beta
         real(c_double), intent(in) :: mass(*)
         integer(c_int64_t), value :: seed
         type(c_ptr), intent(out) :: box
         integer(c_int) :: rc
      end function
      function cam_box_destroy(box) bind(C, name="cam_box_destroy") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: box
         integer(c_int) :: rc
      end function
      function cam_box_set_state(box, q, p) bind(C, name="cam_box_set_state") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: box
         real(c_double), intent(in) :: q(*), p(*)
         integer(c_int) :: rc
      end function
      function cam_box_get_state(box, q, p) bind(C, name="cam_box_get_state") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: box
         real(c_double), intent(out) :: q(*), p(*)
         integer(c_int) :: rc
      end function
      function cam_box_set_nhc(box, nose_q) bind(C, name="cam_box_set_nhc") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: box
         real(c_double), value :: nose_q
         integer(c_int) :: rc
      end function
      function cam_box_set_transrot(box, on) bind(C, name="cam_box_set_transrot") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: box
         integer(c_int), value :: on
         integer(c_int) :: rc
      end function
      function cam_box_run(box, nsteps) bind(C, name="cam_box_run") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: box
         integer(c_int), value :: nsteps
         integer(c_int) :: rc
      end function
      function cam_box_sync(box, ms) bind(C, name="cam_box_sync") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: box
         real(c_double), intent(out) :: ms
         integer(c_int) :: rc
      end function
      function cam_box_energies(box, epot_beads, ekin, espring) bind(C, name="cam_box_energies") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: box
         real(c_double), intent(out) :: epot_beads(*), ekin, espring
         integer(c_int) :: rc
      end function
   end interface

   public :: cam_last_error, cam_egrad_h3, cam_egrad_ch4h, cam_egrad_oh3, cam_egrad_brh2, cam_egrad_o3
   public :: cam_engine_create, cam_engine_destroy, cam_engine_set_state, cam_engine_get_state
   public :: cam_engine_get_forces, cam_engine_draw_momenta, cam_engine_constrain
   public :: cam_engine_begin_children, cam_engine_begin_parent, cam_engine_run, cam_engine_sync, cam_engine_kappa
   public :: cam_engine_set_windows, cam_engine_begin_umbrella, cam_engine_umbrella_stats
   public :: cam_engine_set_umbrella_histogram, cam_engine_umbrella_histogram
   public :: cam_engine_observables, cam_engine_prepare, cam_engine_autotune, cam_engine_last_timing
   public :: cam_batch_create, cam_batch_destroy, cam_batch_engine, cam_batch_autotune, cam_batch_run
   public :: cam_batch_sync, cam_water_create, cam_ff_destroy, cam_ff_set_pair_mode
   public :: cam_qmdff_create, cam_box_create, cam_box_destroy, cam_box_set_state, cam_box_get_state
   public :: cam_box_set_nhc, cam_box_set_transrot, cam_box_run, cam_box_sync, cam_box_energies
   public :: egrad_h3_gpu, egrad_ch4h_gpu, egrad_oh3_gpu, egrad_brh2_gpu, egrad_o3_gpu, cam_error_message

contains

   ! Same argument list as the reference's egrad_h3 / egrad_ch4h, so a host
   ! can swap the call in gradient.f90 (src/gradient.f90:181) unchanged.
   subroutine egrad_h3_gpu(q, natoms, nbeads, v, dvdq, info)
      integer, intent(in) :: natoms, nbeads
      real(c_double), intent(in) :: q(3, natoms, nbeads)
      real(c_double), intent(out) :: v(nbeads), dvdq(3, natoms, nbeads)
      integer, intent(out) :: info
      info = cam_egrad_h3(q, int(natoms, c_int), int(nbeads, c_int), v, dvdq)
   end subroutine egrad_h3_gpu

   subroutine egrad_ch4h_gpu(q, natoms, nbeads, v, dvdq, info)
      integer, intent(in) :: natoms, nbeads
      real(c_double), intent(in) :: q(3, natoms, nbeads)
      real(c_double), intent(out) :: v(nbeads), dvdq(3, natoms, nbeads)
      integer, intent(out) :: info
      info = cam_egrad_ch4h(q, int(natoms, c_int), int(nbeads, c_int), v, dvdq)
   end subroutine egrad_ch4h_gpu

   subroutine egrad_oh3_gpu(q, natoms, nbeads, v, dvdq, info)
      integer, intent(in) :: natoms, nbeads
      real(c_double), intent(in) :: q(3, natoms, nbeads)
      real(c_double), intent(out) :: v(nbeads), dvdq(3, natoms, nbeads)
      integer, intent(out) :: info
      info = cam_egrad_oh3(q, int(natoms, c_int), int(nbeads, c_int), v, dvdq)
   end subroutine egrad_oh3_gpu

   subroutine egrad_brh2_gpu(q, natoms, nbeads, v, dvdq, info)
      integer, intent(in) :: natoms, nbeads
      real(c_double), intent(in) :: q(3, natoms, nbeads)
      real(c_double), intent(out) :: v(nbeads), dvdq(3, natoms, nbeads)
      integer, intent(out) :: info
      info = cam_egrad_brh2(q, int(natoms, c_int), int(nbeads, c_int), v, dvdq)
   end subroutine egrad_brh2_gpu

   subroutine egrad_o3_gpu(q, natoms, nbeads, v, dvdq, info)
      integer, intent(in) :: natoms, nbeads
      real(c_double), intent(in) :: q(3, natoms, nbeads)
      real(c_double), intent(out) :: v(nbeads), dvdq(3, natoms, nbeads)
      integer, intent(out) :: info
      info = cam_egrad_o3(q, int(natoms, c_int), int(nbeads, c_int), v, dvdq)
   end subroutine egrad_o3_gpu

   function cam_error_message() result(msg)
      character(len=:), allocatable :: msg
      character(kind=c_char), pointer :: chars(:)
      type(c_ptr) :: p
      integer :: n
      interface
         function c_strlen(s) bind(C, name="strlen") result(l)
            import :: c_ptr, c_size_t
            type(c_ptr), value :: s
            integer(c_size_t) :: l
         end function
      end interface
      p = cam_last_error()
      n = int(c_strlen(p))
      call c_f_pointer(p, chars, [n])
      allocate(character(len=n) :: msg)
      msg = transfer(chars(1:n), msg)
   end function cam_error_message

end module caracal_amd_mod
