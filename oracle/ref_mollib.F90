! C-callable wrappers around routines of the reference's Src/mollib.F90, compiled from that
! file where it lies into oracle/_ref/ (oracle/Makefile, target "ref").  Test
! infrastructure only: the tests use it to pin the oracle's restatements of
!   Random            Src/mollib.F90:2683-2713
!   CardinalBSpline   Src/mollib.F90:1625-1683
!   ErfLocal / Erf2   Src/mollib.F90:175-274
! against the reference's own compiled code.
module RefMollib
   use iso_c_binding
   use Random_Module, only: am, ix, iy
   use CardinalBSplineModule, only: CardinalBSpline
   implicit none

   interface
      function Random(idum)
         integer(4), intent(inout) :: idum
         real(8) :: Random
      end function Random
      function ErfLocal(x)
         real(8), intent(in) :: x
         real(8) :: ErfLocal
      end function ErfLocal
   end interface

contains

   ! n draws of Random(idum) starting from the generator state (idum, ix, iy, am); the state
   ! after the last draw is written back.
   subroutine ref_random(n, idum, state, amio, out) bind(C, name="ref_random")
      integer(c_int), value :: n
      integer(c_int32_t), intent(inout) :: idum
      integer(c_int32_t), intent(inout) :: state(2)
      real(c_double), intent(inout) :: amio
      real(c_double), intent(out) :: out(n)
      integer(4) :: i, id
      ix = state(1)
      iy = state(2)
      am = amio
      id = idum
      do i = 1, n
         out(i) = Random(id)
      end do
      idum = id
      state(1) = ix
      state(2) = iy
      amio = am
   end subroutine ref_random

   ! CardinalBSpline(order, pos, kmax, values, derivs) at npos positions; values and
   ! derivs are (order, npos), column-major.
   subroutine ref_bspline(order, npos, pos, kmax, values, derivs) bind(C, name="ref_bspline")
      integer(c_int), value :: order, npos
      real(c_double), intent(in) :: pos(npos)
      integer(c_int), intent(out) :: kmax(npos)
      real(c_double), intent(out) :: values(order, npos), derivs(order, npos)
      integer(4) :: i, k
      do i = 1, npos
         call CardinalBSpline(order, pos(i), k, values(:, i), derivs(:, i))
         kmax(i) = k
      end do
   end subroutine ref_bspline

   subroutine ref_erf(n, x, out) bind(C, name="ref_erf")
      integer(c_int), value :: n
      real(c_double), intent(in) :: x(n)
      real(c_double), intent(out) :: out(n)
      integer(4) :: i
      do i = 1, n
         out(i) = ErfLocal(x(i))
      end do
   end subroutine ref_erf

end module RefMollib
