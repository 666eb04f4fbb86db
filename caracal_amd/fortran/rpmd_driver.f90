!
!   rpmd_driver: a minimal dynamic.x-style host over caracal_amd_mod.
!   Reads a run description and the initial ring polymers from a binary
!   stream file, advances them with the GPU engine, and writes the final
!   positions and momenta.
!
!   Input stream (little-endian, no record markers):
!     int32  surface, nbeads, ntraj, ensemble, remove_com, nsteps
!     real64 dt, beta, mass(natoms)
!     real64 q(3,natoms,nbeads,ntraj), p(3,natoms,nbeads,ntraj)
!   Output stream: real64 q(...), p(...) after nsteps steps, then the
!   drop-in surface check V(nbeads*ntraj) from egrad_*_gpu on the input q.
!
!   usage: rpmd_driver input.bin output.bin
!
program rpmd_driver
   use, intrinsic :: iso_c_binding
   use caracal_amd_mod
   implicit none
   character(len=512) :: fin, fout
   integer(c_int32_t) :: hdr(6)
   integer :: natoms, nbeads, ntraj, nsteps, u, info, i
   real(c_double) :: dt, beta
   real(c_double), allocatable :: q(:,:,:,:), p(:,:,:,:), q0(:,:,:,:), bet(:), v(:), g(:,:,:,:)
   type(cam_config) :: cfg
   type(c_ptr) :: eng
   integer(c_int) :: rc

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open(newunit=u, file=trim(fin), access="stream", form="unformatted", status="old")
   read(u) hdr
   select case (hdr(1))
   case (CAM_SURFACE_H3, CAM_SURFACE_BRH2, CAM_SURFACE_O3)
      natoms = 3
   case (CAM_SURFACE_OH3)
      natoms = 4
   case default
      natoms = 6
   end select
   nbeads = hdr(2)
   ntraj = hdr(3)
   nsteps = hdr(6)
   read(u) dt, beta
   cfg%surface = hdr(1)
   cfg%natoms = natoms
   cfg%nbeads = nbeads
   cfg%ntraj = ntraj
   cfg%ensemble = hdr(4)
   cfg%remove_com = hdr(5)
   cfg%dt = dt
   read(u) cfg%mass(1:natoms)
   allocate(q(3,natoms,nbeads,ntraj), p(3,natoms,nbeads,ntraj), bet(ntraj))
   read(u) q, p
   close(u)
   q0 = q
   bet = beta

   rc = cam_engine_create(cfg, eng)
   if (rc /= CAM_OK) call fail("cam_engine_create")
   rc = cam_engine_set_state(eng, q, p, bet)
   if (rc /= CAM_OK) call fail("cam_engine_set_state")
   rc = cam_engine_run(eng, int(nsteps, c_int))
   if (rc /= CAM_OK) call fail("cam_engine_run")
   rc = cam_engine_sync(eng)
   if (rc /= CAM_OK) call fail("cam_engine_sync")
   rc = cam_engine_get_state(eng, q, p)
   if (rc /= CAM_OK) call fail("cam_engine_get_state")
   rc = cam_engine_destroy(eng)

   allocate(v(nbeads*ntraj), g(3,natoms,nbeads,ntraj))
   if (hdr(1) == CAM_SURFACE_H3) then
      call egrad_h3_gpu(q0, natoms, nbeads*ntraj, v, g, info)
   else if (hdr(1) == CAM_SURFACE_OH3) then
      call egrad_oh3_gpu(q0, natoms, nbeads*ntraj, v, g, info)
   else if (hdr(1) == CAM_SURFACE_BRH2) then
      call egrad_brh2_gpu(q0, natoms, nbeads*ntraj, v, g, info)
   else if (hdr(1) == CAM_SURFACE_O3) then
      call egrad_o3_gpu(q0, natoms, nbeads*ntraj, v, g, info)
   else
      call egrad_ch4h_gpu(q0, natoms, nbeads*ntraj, v, g, info)
   end if
   if (info /= 0) call fail("egrad_*_gpu")

   open(newunit=u, file=trim(fout), access="stream", form="unformatted", status="replace")
   write(u) q, p, v
   close(u)
   write(*,'(a,i0,a,i0,a,i0,a)') "rpmd_driver: ", ntraj, " trajectories x ", nbeads, " beads, ", nsteps, " steps done"

contains

   subroutine fail(what)
      character(len=*), intent(in) :: what
      write(*,*) what, " failed: ", cam_error_message()
      stop 1
   end subroutine fail

end program rpmd_driver
