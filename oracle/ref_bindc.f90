! ref_bindc.f90 -- C-callable wrappers around the reference's own compiled routines.
!
! TEST INFRASTRUCTURE ONLY (the oracle's pin, never linked by the product).  oracle/Makefile
! compiles kernel.f90 from the reference tree unchanged, then this file exposes its
! module procedures with bind(C) names so that the tests can compare the C restatement
! (ibx_oracle.c) with the reference's own machine code, bit for bit, and the vendored VFFTPACK
! (libs/fft/fft.f) that pins the oracle's real transforms (fluid_oracle.c).  It also defines the
! COMMON block /MATLAW/ that the reference's Fortran side declares (common.f90:4) and that
! membmodel.c reads (membmodel.c:16-18).
module ref_bindc
  use iso_c_binding
  use mod_kernel
  implicit none
  real(c_double) :: hf_gscell
  common /matlaw/ hf_gscell
contains
  function ref_kernel(r) bind(C, name="ref_kernel") result(v)
    real(c_double), value :: r
    real(c_double) :: v
    real :: rr
    rr = r
    v = kernel(rr)
  end function ref_kernel

  function ref_kerneltest() bind(C, name="ref_kerneltest") result(v)
    real(c_double) :: v
    real :: s
    call kerneltest(s)
    v = s
  end function ref_kerneltest

  subroutine ref_set_hf_gscell(h) bind(C, name="ref_set_hf_gscell")
    real(c_double), value :: h
    hf_gscell = h
  end subroutine ref_set_hf_gscell

  ! VFFTPACK as the reference calls it from solver1d (solver.f90:61-66,139-145): one sequence
  ! (M = 1, MDIMR = 1) of length n, initialised by vrffti (initsolver.f90:102-105)
  subroutine ref_vrfft(n, x, inverse) bind(C, name="ref_vrfft")
    integer(c_int), value :: n, inverse
    real(c_double), intent(inout) :: x(n)
    real :: r(1, n), rt(1, n), wsave(n + 15)
    external vrffti, vrfftf, vrfftb
    r(1, :) = x
    call vrffti(n, wsave)
    if (inverse /= 0) then
      call vrfftb(1, n, r, rt, 1, wsave)
    else
      call vrfftf(1, n, r, rt, 1, wsave)
    end if
    x = r(1, :)
  end subroutine ref_vrfft
end module ref_bindc
