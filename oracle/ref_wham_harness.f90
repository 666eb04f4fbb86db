! oracle/ref_wham_harness.f90 -- TEST INFRASTRUCTURE ONLY.
! A driver program around the reference's own WHAM routine (src/wham.f90),
! compiled from the reference's sources by Makefile.ref.  It sets what
! calc_rate sets before the call (the temperature in general, the window
! force constants in evb_mod, src/calc_rate.f90:1793-1803) and calls wham
! unchanged.  wham itself writes its progress to unit 15 and pmf_wham.dat in
! the working directory.
!
! Usage: ref_wham_harness <input> <output>
! Input (stream): int32 nbins, n_all; real64 xi_min, xi_max, kelvin;
!   real64 xi_wins(n_all), k_force(n_all), hist(n_all, nbins) (column-major:
!   hist(i, j), window i, bin j).
! Output (stream): real64 pmf(nbins) as wham returns it.
program ref_wham_harness
   use general
   use evb_mod
   implicit none
   character(len=512) :: fin, fout
   integer(4) :: nb4, na4
   integer :: nbins, nall
   real(kind=8) :: xmin, xmax, bsize
   real(kind=8), allocatable :: xw(:), hist(:, :), pmf(:)

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open(unit=71, file=trim(fin), access="stream", form="unformatted", status="old")
   read(71) nb4, na4
   nbins = nb4
   nall = na4
   read(71) xmin, xmax, kelvin
   allocate(xw(nall), hist(nall, nbins), pmf(nbins), k_force(nall))
   read(71) xw, k_force, hist
   close(71)
   bsize = (xmax - xmin) / (nbins - 1)
   call wham(nbins, nall, xmin, xmax, bsize, xw, hist, pmf)
   open(unit=72, file=trim(fout), access="stream", form="unformatted", status="replace")
   write(72) pmf
   close(72)
end program ref_wham_harness
