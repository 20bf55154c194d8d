! C-callable entry points into the reference's primitive-integral module
! and its error function.
!
! Checker only: this file is compiled together with the reference's own
! sources (src/kinds.f90, src/constants.f90, src/imago/gaussIntegrals.f90,
! src/imago/mathSubs.f90) into oracle/_ref/libgauss_ref.so, which tests/golden/make_golden.py uses
! to write golden vectors.  Nothing in the product links it.
!
! Each routine zeroes sh before the call: the reference routines only write
! the leading lmAlphaIndex(a1) x lmAlphaIndex(a2) block of sh(16,16)
! (gaussIntegrals.f90:26-4420 and the other integral routines).
module imago_ref_bindc
   use iso_c_binding
   use O_Kinds
   use O_GaussianIntegrals
   use O_MathSubs, only: error_fn
   implicit none
contains

   subroutine ref_overlap2c(a1, a2, A, B, sw, sh) bind(C, name="ref_overlap2c")
      real(c_double), value :: a1, a2
      real(c_double), intent(in) :: A(3), B(3)
      integer(c_int), value :: sw
      real(c_double), intent(out) :: sh(16,16)
      sh = 0.0_double
      call overlap2CIntg(a1, a2, A, B, int(sw), sh)
   end subroutine ref_overlap2c

   subroutine ref_kinetic2c(a1, a2, A, B, sw, sh) bind(C, name="ref_kinetic2c")
      real(c_double), value :: a1, a2
      real(c_double), intent(in) :: A(3), B(3)
      integer(c_int), value :: sw
      real(c_double), intent(out) :: sh(16,16)
      sh = 0.0_double
      call kinetic2CIntg(a1, a2, A, B, int(sw), sh)
   end subroutine ref_kinetic2c

   subroutine ref_electron3c(a1, a2, a3, A, B, C, sw, sh) &
         & bind(C, name="ref_electron3c")
      real(c_double), value :: a1, a2, a3
      real(c_double), intent(in) :: A(3), B(3), C(3)
      integer(c_int), value :: sw
      real(c_double), intent(out) :: sh(16,16)
      sh = 0.0_double
      call electron3CIntg(a1, a2, a3, A, B, C, int(sw), sh)
   end subroutine ref_electron3c

   subroutine ref_nuclear3c(a1, a2, a3, A, B, C, sw, sh) &
         & bind(C, name="ref_nuclear3c")
      real(c_double), value :: a1, a2, a3
      real(c_double), intent(in) :: A(3), B(3), C(3)
      integer(c_int), value :: sw
      real(c_double), intent(out) :: sh(16,16)
      sh = 0.0_double
      call nuclear3CIntg(a1, a2, a3, A, B, C, int(sw), sh)
   end subroutine ref_nuclear3c

   subroutine ref_boys(T, F) bind(C, name="ref_boys")
      real(c_double), value :: T
      real(c_double), intent(out) :: F(7)
      call boys(T, F)
   end subroutine ref_boys

   function ref_error_fn(x) bind(C, name="ref_error_fn") result(v)
      real(c_double), value :: x
      real(c_double) :: v
      real(double) :: xx
      xx = x
      v = error_fn(xx)
   end function ref_error_fn

end module imago_ref_bindc
