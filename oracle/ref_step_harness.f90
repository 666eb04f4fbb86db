! oracle/ref_step_harness.f90 -- TEST INFRASTRUCTURE ONLY.
! A driver program around the reference's own RPMD step (src/verlet.f90) with
! its normal-mode transforms (src/rfft.f90, src/irfft.f90 over the FFTW3 API
! of MKL), transrot, get_centroid, umbrella mode 1 / calc_xi / calc_com and
! the analytic surfaces, all compiled from the reference's sources by
! Makefile.ref.  It sets the module state the reference's input readers set
! (src/dynamic.f90, src/calc_rate_read.f90), calls the reference's mdinit
! (start forces, the umbrella bias of a window start, the Nose-Hoover chain
! masses) with the momenta it is given, and calls verlet unchanged.
!
! Usage: ref_step_harness <input> <output>
! Input (stream, native endianness):
!   int32  natoms, nbeads, nsteps, constrain (-1 dynamic.x, 0 umbrella
!          window, 2 recrossing child), surface (1 H3, 2 CH4H, 3 OH3, 4 QMDFF,
!          5 BrH2, 6 O3),
!          every (write a frame after every 'every' steps; 0 = timing only),
!          thermostat (0 none, 1 Andersen, 2 Nose-Hoover chain), andersen_step
!   real64 dt, beta, mass(natoms), q(3,natoms,nbeads), p(3,natoms,nbeads)
!   thermostat 2:  real64 kelvin, nose_q
!   surface 4:     int32 length, then the QMDFF file name; int32 periodic,
!                  zahn; real64 box, coul_cut, vdw_cut (angstrom, key-file units)
!   constrain >= 0 adds the BIMOLEC reaction coordinate:
!   int32  n_form, n_break, bond_form(2,n_form), bond_break(2,n_break)
!   real64 form_ref(n_form), break_ref(n_break)
!   int32  n_reac, reac_natoms(n_reac), reac_atoms(maxr,n_reac) (maxr = max count)
!   real64 r_inf, xi_ideal
!   constrain = 0: real64 k_force
! Output (stream): per frame  real64 q(3,natoms,nbeads), p(3,natoms,nbeads),
!   epot, xi_real; the step loop's wall time is printed to stdout.
program ref_step_harness
   use general
   use evb_mod
   use qmdff
   use pbc_mod
   implicit none
   include 'mpif.h'
   integer :: ierr
   integer :: nat, nb, nsteps, cmode, surf, every, ithermo, astep, i, k, istep, nf, nbk, nr, maxr
   integer :: rank, round
   integer, allocatable :: bf(:, :), bb(:, :), rn(:), ra(:, :)
   real(kind=8) :: dt, beta_in, epot, ekin, afm_force, xi_ideal, xi_real, kf
   real(kind=8), allocatable :: derivs(:, :, :), dxi_act(:, :), p_in(:, :, :)
   integer(kind=8) :: c0, c1, crate
   character(len=512) :: fin, fout
   character(len=70) :: ff1, ff2, ff3
   integer :: plen, iper, izahn
   real(kind=8) :: boxa, ccut, vcut
   logical :: analyze

   ! dynamic.x runs under MPI (src/dynamic.f90); verlet's dynamic.x branch
   ! broadcasts the state from rank 0 (src/verlet.f90:767-771)
   call mpi_init(ierr)
   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open (unit=71, file=trim(fin), access='stream', form='unformatted', status='old')
   read (71) nat, nb, nsteps, cmode, surf, every, ithermo, astep
   read (71) dt, beta_in
   natoms = nat
   nbeads = nb
   read (71) mass(1:nat)
   allocate (q_i(3, nat, nb), p_i(3, nat, nb), p_in(3, nat, nb), derivs(3, nat, nb), dxi_act(3, nat))
   read (71) q_i
   read (71) p_in
   kelvin = 1.d0 / (0.316679d-5 * beta_in)
   nose_q = 0.d0
   if (ithermo == 2) read (71) kelvin, nose_q
   if (surf == 4) then
      read (71) plen
      ff1 = ""
      read (71) ff1(1:plen)
      read (71) iper, izahn
      read (71) boxa, ccut, vcut
   end if
   xi_ideal = 0.d0
   if (cmode >= 0) then
      read (71) nf, nbk
      form_num = nf
      break_num = nbk
      allocate (bf(2, nf), bb(2, nbk), bond_form(nf, 2), bond_break(nbk, 2), form_ref(nf), break_ref(nbk))
      read (71) bf
      read (71) bb
      bond_form(:, 1) = bf(1, :)
      bond_form(:, 2) = bf(2, :)
      bond_break(:, 1) = bb(1, :)
      bond_break(:, 2) = bb(2, :)
      read (71) form_ref
      read (71) break_ref
      read (71) nr
      sum_reacs = nr
      allocate (rn(nr), n_reac(nr), mass_reac(nr))
      read (71) rn
      maxr = maxval(rn)
      allocate (ra(maxr, nr), at_reac(nr, maxr))
      read (71) ra
      n_reac = rn
      at_reac = 0
      do k = 1, nr
         at_reac(k, 1:rn(k)) = ra(1:rn(k), k)
         mass_reac(k) = 0.d0
         do i = 1, rn(k)
            mass_reac(k) = mass_reac(k) + mass(ra(i, k))
         end do
      end do
      read (71) R_inf, xi_ideal
      umbr_type = "BIMOLEC"
   end if
   if (cmode == 0) then
      read (71) kf
      allocate (k_force(1))
      k_force(1) = kf
      um_window_act = 1
   end if
   close (71)

   ! module state of a dynamic.x / calc_rate run with no extras
   beta = beta_in
   pot_ana = .true.
   select case (surf)
   case (1)
      pot_type = "h3"
   case (2)
      pot_type = "ch4h"
      call initialize_ch4h()
   case (3)
      pot_type = "oh3"
      call initialize_oh3()
   case (5)
      pot_type = "brh2"
      call initialize_brh2()
   case (6)
      pot_type = "o3"
   case (4)
      ! one QMDFF (src/read_pes.f90:1745-1919, :2729-2736); gradient then
      ! takes its ff_eg / ff_nonb / ff_hb branch (src/gradient.f90:341-355)
      pot_ana = .false.
      pot_type = "qmdff"
      ewald = .false.
      zahn = izahn /= 0
      coul_cut = ccut
      vdw_cut = vcut
      eval_cutoff = .false.
      ff2 = "none"
      ff3 = "none"
      call prepare(ff1, ff2, ff3, 1)
      nqmdff = 1
      treq = .false.
      E_zero1 = 0.d0
      water_spc = .false.
      gfn_xtb = .false.
      pgfn_ff = .false.
      orca = .false.
      aenet_ann = .false.
      mace_ase = .false.
      uma_ase = .false.
      call_ext = .false.
      call_cust = .false.
      periodic = iper /= 0
      if (periodic) then
         boxlen_x = boxa / bohr
         boxlen_y = boxa / bohr
         boxlen_z = boxa / bohr
         call set_periodic(periodic, boxlen_x, boxlen_y, boxlen_z, ewald, zahn, ewald_brute, coul_cut, vdw_cut)
      end if
   end select
   use_calc_rate = cmode >= 0
   use_stick_coeff = .false.
   use_rpmd = .true.
   thermostat = ithermo
   andersen_step = max(astep, 1)
   nve = .true.
   read_vel = .false.
   fix_atoms = .false.
   barostat = 0
   npt = .false.
   if (surf /= 4) periodic = .false.
   box_walls = .false.
   mirrors = .false.
   calc_ekin = .false.
   energysplit = .false.
   verbose = .false.
   print_gen = .false.
   print_cross = .false.
   print_train = .false.
   output_sparse = .false.
   add_force = .false.
   afm_run = .false.
   bias_list = .false.
   coord_vasp = .false.
   calc_vir = .false.
   nat6 = 0
   iwrite = huge(1)
   print_gen_freq = huge(1)
   print_cross_freq = huge(1)
   allocate (at_move(nat))
   at_move = .true.
   rank = 0
   round = 0
   analyze = .false.
   afm_force = 0.d0

   ! mdinit: start forces of every bead, the umbrella bias of a window start
   ! (bias_mode 2, as src/calc_rate.f90:1554 calls it), the chain masses of
   ! the Nose-Hoover thermostat; its momentum draw is replaced by the given
   ! momenta (nve makes it set p = 0, or the chain's andersen call draws them)
   if (cmode == 0) then
      call mdinit(derivs, xi_ideal, dxi_act, 2, rank)
   else
      call mdinit(derivs, xi_ideal, dxi_act, 0, rank)
   end if
   p_i = p_in

   open (unit=72, file=trim(fout), access='stream', form='unformatted', status='replace')
   call system_clock(c0, crate)
   do istep = 1, nsteps
      call verlet(istep, dt, derivs, epot, ekin, afm_force, xi_ideal, xi_real, dxi_act, round, cmode, analyze, rank)
      if (every > 0) then
         if (mod(istep, every) == 0) write (72) q_i, p_i, epot, xi_real
      end if
   end do
   call system_clock(c1)
   close (72)
   write (*, '(a,es24.16)') 'seconds ', dble(c1 - c0) / dble(crate)
   call mpi_finalize(ierr)
end program ref_step_harness
