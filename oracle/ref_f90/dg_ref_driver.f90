! dg_ref_driver.f90 -- harness around the reference's Fortran physics module.
!
! TEST INFRASTRUCTURE ONLY.  Linked against src/f90/single/physics.f90 of the
! reference (compiled where it lies by oracle/Makefile), it runs the reference
! time-loop body (the call sequence of src/f90/single/euler_cfd.f90:107-156)
! on an initial state read from a file, without the HDF5 output, and writes
! the conserved fields and the per-step dt.  Used only to write golden
! vectors (tests/golden/) and as the reference CPU timing.
!
! Input file (stream, little endian):
!   int32 nxi, nyi, nzi, nsteps, bcs(6), nthreads; real(8) ds
!   real(8) rho, vx, vy, vz, p on the ghosted grid, Fortran order (x fastest)
! Output file:
!   real(8) dt(nsteps), wall_seconds
!   real(8) mass, momentum_x, momentum_y, momentum_z, energy (ghosted)
program dg_ref_driver
   use types_and_kinds
   use global
   use physics
   use omp_lib
   implicit none
   character(len=4096) :: fin, fout
   integer(4) :: hdr(4), bcs(6), nthr, uin, uout
   integer(ik) :: nx, ny, nz, nsteps, step
   integer(ik), parameter :: ng = 2
   integer(ik) :: shiftx(3), shifty(3), shiftz(3)
   real(8) :: ds8, t0, t1
   real(rk) :: ds, dt
   real(8), allocatable :: buf(:, :, :), dts(:)
   real(rk), allocatable, dimension(:, :, :) :: &
      rho, vx, vy, vz, p, mass, mx, my, mz, en, temp, &
      fmx, fmy, fmz, fxx, fxy, fxz, fyx, fyy, fyz, fzx, fzy, fzz, fex, fey, fez, &
      drx, dry, drz, dxx, dxy, dxz, dyx, dyy, dyz, dzx, dzy, dzz, dpx, dpy, dpz, &
      rxt, vxt, vyt, vzt, pxt

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open (newunit=uin, file=trim(fin), access='stream', form='unformatted', status='old')
   read (uin) hdr, bcs, nthr, ds8
   if (nthr > 0) call omp_set_num_threads(nthr)
   nx = hdr(1) + 2*ng
   ny = hdr(2) + 2*ng
   nz = hdr(3) + 2*ng
   nsteps = hdr(4)
   ds = real(ds8, rk)

   allocate (buf(nx, ny, nz), dts(max(nsteps, 1)))
   call alloc0(rho); call alloc0(vx); call alloc0(vy); call alloc0(vz); call alloc0(p)
   call alloc0(mass); call alloc0(mx); call alloc0(my); call alloc0(mz); call alloc0(en)
   call alloc0(temp)
   call alloc0(fmx); call alloc0(fmy); call alloc0(fmz)
   call alloc0(fxx); call alloc0(fxy); call alloc0(fxz)
   call alloc0(fyx); call alloc0(fyy); call alloc0(fyz)
   call alloc0(fzx); call alloc0(fzy); call alloc0(fzz)
   call alloc0(fex); call alloc0(fey); call alloc0(fez)
   call alloc0(drx); call alloc0(dry); call alloc0(drz)
   call alloc0(dxx); call alloc0(dxy); call alloc0(dxz)
   call alloc0(dyx); call alloc0(dyy); call alloc0(dyz)
   call alloc0(dzx); call alloc0(dzy); call alloc0(dzz)
   call alloc0(dpx); call alloc0(dpy); call alloc0(dpz)
   call alloc0(rxt); call alloc0(vxt); call alloc0(vyt); call alloc0(vzt); call alloc0(pxt)

   read (uin) buf; rho = real(buf, rk)
   read (uin) buf; vx = real(buf, rk)
   read (uin) buf; vy = real(buf, rk)
   read (uin) buf; vz = real(buf, rk)
   read (uin) buf; p = real(buf, rk)
   close (uin)

   shiftx = 0; shifty = 0; shiftz = 0
   shiftx(1) = 1; shifty(2) = 1; shiftz(3) = 1

   ! set-up as euler_cfd.f90:100-101
   call update_primitive_ghosts(rho, vx, vy, vz, p, nx, ny, nz, ng, bcs)
   call conservative(mass, mx, my, mz, en, rho, p, vx, vy, vz, temp, ds)

   t0 = omp_get_wtime()
   do step = 1, nsteps
      ! loop body as euler_cfd.f90:107-156
      call apply_boundary_conditions(rho, vx, vy, vz, p, nx, ny, nz, ng, bcs)
      call update_primitive_ghosts(rho, vx, vy, vz, p, nx, ny, nz, ng, bcs)
      call primitive(mass, mx, my, mz, en, rho, p, vx, vy, vz, temp, ds)
      call update_primitive_ghosts(rho, vx, vy, vz, p, nx, ny, nz, ng, bcs)
      dt = compute_timestep(ds, vx, vy, vz, p, rho)
      dts(step) = real(dt, 8)
      call calculate_gradients(rho, drx, dry, drz, nx, ny, nz, ng, ds)
      call calculate_gradients(vx, dxx, dxy, dxz, nx, ny, nz, ng, ds)
      call calculate_gradients(vy, dyx, dyy, dyz, nx, ny, nz, ng, ds)
      call calculate_gradients(vz, dzx, dzy, dzz, nx, ny, nz, ng, ds)
      call calculate_gradients(p, dpx, dpy, dpz, nx, ny, nz, ng, ds)
      call update_gradient_ghost(drx, dry, drz, dxx, dxy, dxz, dyx, dyy, dyz, dzx, dzy, dzz, &
                                 dpx, dpy, dpz, nx, ny, nz, ng, bcs)
      call extrapolate_primitves(rho, vx, vy, vz, p, drx, dry, drz, dxx, dxy, dxz, dyx, dyy, dyz, &
                                 dzx, dzy, dzz, dpx, dpy, dpz, rxt, vxt, vyt, vzt, pxt, dt)
      call apply_boundary_conditions(rxt, vxt, vyt, vzt, pxt, nx, ny, nz, ng, bcs)
      call reconstructflux(fmx, fxx, fyx, fzx, fex, drx, dxx, dyx, dzx, dpx, &
                           rxt, vxt, vyt, vzt, pxt, nx, ny, nz, ng, ds, shiftx)
      call reconstructflux(fmy, fyy, fxy, fzy, fey, dry, dyy, dxy, dzy, dpy, &
                           rxt, vyt, vxt, vzt, pxt, nx, ny, nz, ng, ds, shifty)
      call reconstructflux(fmz, fzz, fxz, fyz, fez, drz, dzz, dxz, dyz, dpz, &
                           rxt, vzt, vxt, vyt, pxt, nx, ny, nz, ng, ds, shiftz)
      call update_flux_ghost(fmx, fmy, fmz, fxx, fyx, fzx, fxy, fyy, fzy, fxz, fyz, fzz, &
                             fex, fey, fez, nx, ny, nz, ng, bcs)
      call addfluxes(fmx, fmy, fmz, fxx, fxy, fxz, fyx, fyy, fyz, fzx, fzy, fzz, fex, fey, fez, &
                     mass, mx, my, mz, en, nx, ny, nz, ng, dt, ds)
   end do
   t1 = omp_get_wtime()

   open (newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace')
   write (uout) dts(1:max(nsteps, 1)), t1 - t0
   buf = real(mass, 8); write (uout) buf
   buf = real(mx, 8); write (uout) buf
   buf = real(my, 8); write (uout) buf
   buf = real(mz, 8); write (uout) buf
   buf = real(en, 8); write (uout) buf
   close (uout)

contains
   subroutine alloc0(a)
      real(rk), allocatable, intent(inout) :: a(:, :, :)
      allocate (a(nx, ny, nz))
      a = 0.0_rk
   end subroutine alloc0
end program dg_ref_driver
