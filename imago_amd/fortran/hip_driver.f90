! hip_driver.f90 - a minimal Fortran host that calls the library through
! O_ImagoHIP the way the reference's secularEqnSCF calls solveZHEGV: a
! 2x2 Hermitian generalized problem with known eigenvalues.
! Build: amdflang imago_hip.f90 hip_driver.f90 -L.. -limago_hip
program hip_driver
   use iso_c_binding
   use O_ImagoHIP
   implicit none
   complex(c_double_complex) :: A(2,2), B(2,2)
   real(c_double) :: w(2)
   ! H = [[2, i],[-i, 2]], S = I: eigenvalues 1 and 3
   A = reshape([(2.0d0,0.0d0), (0.0d0,-1.0d0), (0.0d0,1.0d0), (2.0d0,0.0d0)], [2,2])
   B = reshape([(1.0d0,0.0d0), (0.0d0,0.0d0), (0.0d0,0.0d0), (1.0d0,0.0d0)], [2,2])
   call solveZHEGV_hip(2, 2, A, B, w)
   write (*,"(a,2f12.8)") 'eigenvalues', w
   if (abs(w(1) - 1.0d0) > 1.0d-12 .or. abs(w(2) - 3.0d0) > 1.0d-12) stop 1
end program hip_driver
