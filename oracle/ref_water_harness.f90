! oracle/ref_water_harness.f90 -- TEST INFRASTRUCTURE ONLY.
! A C-callable driver around the reference's own water surface
! (src/egrad_water.f90, src/water_init.f90), compiled from the reference's
! sources by Makefile.ref.  It sets the module state the reference's input
! reader would set (atom names O, H, H, ...; periodic box; Zahn parameters)
! and calls water_init and egrad_water unchanged.
subroutine ref_water_eval(nat, xw, gw, ew, iper, box, izahn, za, zp, cut) bind(C, name="ref_water_eval")
   use iso_c_binding
   use general
   use evb_mod
   use pbc_mod
   use qmdff
   implicit none
   integer(c_int), value :: nat, iper, izahn
   real(c_double), value :: box, za, zp, cut
   real(c_double), intent(in) :: xw(3, nat)
   real(c_double), intent(out) :: gw(3, nat), ew
   integer :: i
   natoms = nat
   do i = 1, nat
      if (mod(i - 1, 3) == 0) then
         name(i) = "O"
      else
         name(i) = "H"
      end if
   end do
   periodic = iper /= 0
   boxlen_x = box
   boxlen_y = box
   boxlen_z = box
   boxlen_x2 = 0.5d0 * box
   boxlen_y2 = 0.5d0 * box
   boxlen_z2 = 0.5d0 * box
   zahn = izahn /= 0
   zahn_a = za
   zahn_par = zp
   coul_cut = cut
   calc_vir = .false.
   if (allocated(water_act)) deallocate(water_act)
   if (allocated(water_pars)) deallocate(water_pars)
   if (allocated(q)) deallocate(q)
   call water_init
   call egrad_water(xw, gw, ew)
end subroutine ref_water_eval
