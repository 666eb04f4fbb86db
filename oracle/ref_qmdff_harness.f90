! oracle/ref_qmdff_harness.f90 -- TEST INFRASTRUCTURE ONLY.
! A driver program around the reference's own QMDFF force field, compiled
! from the reference's sources by Makefile.ref: prepare (rdsolvff, setnonb,
! copyc6, ncoord_qmdff, getc6), set_periodic, and the per-structure call
! sequence of gradient (src/gradient.f90:341-348): ff_eg, ff_nonb, ff_hb.
! It sets what the reference's key-file reader sets for a QMDFF (src/read_pes.f90:1745-1919,
! :2729-2736) and calls the routines unchanged.
!
! Usage: ref_qmdff_harness <file.qmdff> <input> <output> [<dump>]
! Input (stream): int32 periodic, zahn, nconf; real64 box, coul_cut, vdw_cut
!   (angstrom, as the key file gives them); real64 xyz(3,n,nconf) in bohr.
! Output (stream): per structure real64 e_eg, e_nonb, e_hb (energies after
!   ff_eg, after ff_nonb and after ff_hb, cumulative as gradient sums them)
!   and g(3,n).
! Dump (stream, optional): the parameters prepare leaves in the qmdff
!   module and the cutoffs of set_periodic -- int32 n, nbond, nangl, ntors,
!   nhb, nnci, nmols, at(n); real64 molnum(n), q(n); int32 bond(2,nbond);
!   real64 vbond(3,nbond); int32 angl(3,nangl); real64 vangl(2,nangl);
!   int32 tors(6,ntors); real64 vtors(14,ntors); int32 hb(3,nhb); real64
!   vhb(2,nhb); int32 nci(3,nnci); real64 c6xy(n,n), r094_mod(94,94),
!   sr42(94,94), zab(94,94), r0ab(94,94), eps1(6), eps2(6), rad(94),
!   scalehb_glob(94), scalexb_glob(94), coul_cut, vdw_cut, cut_low, zahn_a,
!   zahn_par, boxlen_x, boxlen_y, boxlen_z; int32 periodic, zahn.
program ref_qmdff_harness
   use general
   use evb_mod
   use qmdff
   use pbc_mod
   implicit none
   character(len=512) :: fff, fin, fout, fdump
   character(len=70) :: f1, f2, f3
   integer :: iper, izahn, nconf, k, nat, i
   real(kind=8) :: box, ccut, vcut, e, e1, e2
   real(kind=8), allocatable :: xin(:, :, :), gg(:, :)

   call get_command_argument(1, fff)
   call get_command_argument(2, fin)
   call get_command_argument(3, fout)
   fdump = ""
   if (command_argument_count() >= 4) call get_command_argument(4, fdump)
   open (unit=71, file=trim(fin), access='stream', form='unformatted', status='old')
   read (71) iper, izahn, nconf
   read (71) box, ccut, vcut

   ! the QMDFF block of the key file: Coulomb method, cutoffs (angstrom)
   ewald = .false.
   zahn = izahn /= 0
   coul_cut = ccut
   vdw_cut = vcut
   eval_cutoff = .false.
   calc_vir = .false.
   f1 = fff(1:70)
   f2 = "none"
   f3 = "none"
   call prepare(f1, f2, f3, 1)
   nat = natoms
   ! the periodic box (src/read_pes.f90:2729-2736)
   periodic = iper /= 0
   if (periodic) then
      boxlen_x = box / bohr
      boxlen_y = box / bohr
      boxlen_z = box / bohr
      call set_periodic(periodic, boxlen_x, boxlen_y, boxlen_z, ewald, zahn, ewald_brute, coul_cut, vdw_cut)
   end if

   allocate (xin(3, nat, nconf), gg(3, nat))
   read (71) xin
   close (71)
   open (unit=72, file=trim(fout), access='stream', form='unformatted', status='replace')
   do k = 1, nconf
      call ff_eg(nat, at, xin(:, :, k), e, gg)
      e1 = e
      call ff_nonb(nat, at, xin(:, :, k), q, r0ab, zab, r094_mod, sr42, c6xy, e, gg)
      e2 = e
      call ff_hb(nat, at, xin(:, :, k), e, gg)
      write (72) e1, e2, e, gg
   end do
   close (72)

   if (len_trim(fdump) > 0) then
      open (unit=73, file=trim(fdump), access='stream', form='unformatted', status='replace')
      write (73) nat, nbond, nangl, ntors, nhb, nnci, nmols, at(1:nat)
      write (73) molnum(1:nat), q(1:nat)
      write (73) bond(:, 1:nbond), vbond(:, 1:nbond)
      write (73) angl(:, 1:nangl), vangl(:, 1:nangl)
      write (73) tors(:, 1:ntors), vtors(:, 1:ntors)
      write (73) hb(:, 1:nhb), vhb(:, 1:nhb)
      write (73) nci(:, 1:nnci)
      write (73) c6xy, r094_mod, sr42, zab, r0ab, eps1, eps2, rad, scalehb_glob, scalexb_glob
      write (73) coul_cut, vdw_cut, cut_low, zahn_a, zahn_par, boxlen_x, boxlen_y, boxlen_z
      i = 0
      if (periodic) i = 1
      write (73) i
      i = 0
      if (zahn) i = 1
      write (73) i
      close (73)
   end if
end program ref_qmdff_harness
