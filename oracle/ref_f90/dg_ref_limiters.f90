! dg_ref_limiters.f90 -- evaluates the slope limiters of the reference's
! Fortran physics module (minmod, maxmod and vanleerlimiter of
! src/f90/single/physics.f90, compiled where it lies by oracle/Makefile).
!
! TEST INFRASTRUCTURE ONLY.  The reference's step never calls these functions;
! they pin the limited slopes of the oracle (tests/golden/limiters.npz).
!
! Input file (stream): int32 n; real(8) a(n), b(n)
! Output file (stream): real(8) minmod(n), maxmod(n), vanleerlimiter(n),
! each evaluated in the module's real kind.
program dg_ref_limiters
   use types_and_kinds
   use physics
   implicit none
   character(len=4096) :: fin, fout
   integer(4) :: n, i, uin, uout
   real(8), allocatable :: a(:), b(:), r(:, :)
   real(rk) :: x, y

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open (newunit=uin, file=trim(fin), access='stream', form='unformatted', status='old')
   read (uin) n
   allocate (a(n), b(n), r(n, 3))
   read (uin) a
   read (uin) b
   close (uin)
   do i = 1, n
      x = real(a(i), rk)
      y = real(b(i), rk)
      r(i, 1) = real(minmod(x, y), 8)
      r(i, 2) = real(maxmod(x, y), 8)
      r(i, 3) = real(vanleerlimiter(x, y), 8)
   end do
   open (newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace')
   write (uout) r(:, 1)
   write (uout) r(:, 2)
   write (uout) r(:, 3)
   close (uout)
end program dg_ref_limiters
