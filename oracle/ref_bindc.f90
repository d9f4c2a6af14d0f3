! ref_bindc.f90 -- C-callable wrappers around the reference's own compiled routines.
!
! TEST INFRASTRUCTURE ONLY (the oracle's pin, never linked by the product).  oracle/Makefile
! compiles kernel.f90 from the reference tree unchanged, then this file exposes its
! module procedures with bind(C) names so that the tests can compare the C restatement
! (ibx_oracle.c) with the reference's own machine code, bit for bit.  It also defines the
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
end module ref_bindc
