! ibx_cabi.f90 -- iso_c_binding interface to libibx (declarations in include/ibx.h).
!
! Compile with the reference's real kind settings (-fdefault-real-8 -fdefault-double-8, the
! flang spelling of -autodouble -fpconstant, reference Makefile:7); every dummy here is an
! explicit c_double / c_int64_t so the binding does not depend on them.
module ibx_cabi
  use iso_c_binding
  implicit none
  private
  public :: ibx_grid, ibx_create, ibx_destroy, ibx_last_error, ibx_eulr2lagr, ibx_lagr2eulr, &
            ibx_lagr2eulrind, ibx_update_pos, ibx_kernel, ibx_check

  ! include/ibx.h: typedef struct ibx_grid
  type, bind(C) :: ibx_grid
    integer(c_int32_t) :: itot, jtot, ktot, reserved
    real(c_double) :: lx, ly, lz
    real(c_double) :: dxi, dyi, dzi
    real(c_double) :: dx, dy, dz
    real(c_double) :: boundleft, boundfront
  end type ibx_grid

  interface
    integer(c_int) function ibx_create(ctx, grid, device) bind(C, name="ibx_create")
      import :: c_int, c_ptr, ibx_grid
      type(c_ptr), intent(out) :: ctx
      type(ibx_grid), intent(in) :: grid
      integer(c_int), value :: device
    end function ibx_create

    subroutine ibx_destroy(ctx) bind(C, name="ibx_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine ibx_destroy

    type(c_ptr) function ibx_last_error(ctx) bind(C, name="ibx_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function ibx_last_error

    real(c_double) function ibx_kernel(r) bind(C, name="ibx_kernel")
      import :: c_double
      real(c_double), value :: r
    end function ibx_kernel

    integer(c_int) function ibx_eulr2lagr(ctx, u, v, w, ncap, cap_off, x, y, z, ul, vl, wl) &
        bind(C, name="ibx_eulr2lagr")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: u(*), v(*), w(*)
      integer(c_int64_t), value :: ncap
      integer(c_int64_t), intent(in) :: cap_off(*)
      real(c_double), intent(in) :: x(*), y(*), z(*)
      real(c_double), intent(out) :: ul(*), vl(*), wl(*)
    end function ibx_eulr2lagr

    integer(c_int) function ibx_lagr2eulr(ctx, ncap, cap_off, x, y, z, fx, fy, fz, area, &
                                          forcex, forcey, forcez) bind(C, name="ibx_lagr2eulr")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: ncap
      integer(c_int64_t), intent(in) :: cap_off(*)
      real(c_double), intent(in) :: x(*), y(*), z(*), fx(*), fy(*), fz(*), area(*)
      real(c_double), intent(out) :: forcex(*), forcey(*), forcez(*)
    end function ibx_lagr2eulr

    integer(c_int) function ibx_lagr2eulrind(ctx, ncap, cap_off, x, y, z, xn, yn, zn, area, &
                                             gx, gy, gz) bind(C, name="ibx_lagr2eulrind")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: ncap
      integer(c_int64_t), intent(in) :: cap_off(*)
      real(c_double), intent(in) :: x(*), y(*), z(*), xn(*), yn(*), zn(*), area(*)
      real(c_double), intent(out) :: gx(*), gy(*), gz(*)
    end function ibx_lagr2eulrind

    integer(c_int) function ibx_update_pos(ctx, ncap, cap_off, dt, x, y, z, xold, yold, ul, vl, &
                                           wl, centroid, umean, nclamp) &
        bind(C, name="ibx_update_pos")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: ncap
      integer(c_int64_t), intent(in) :: cap_off(*)
      real(c_double), value :: dt
      real(c_double), intent(inout) :: x(*), y(*), z(*), xold(*), yold(*)
      real(c_double), intent(in) :: ul(*), vl(*), wl(*)
      real(c_double), intent(out) :: centroid(*), umean(*)
      integer(c_int64_t), intent(out) :: nclamp
    end function ibx_update_pos
  end interface

contains

  ! Stop with the library's message, in the reference's fatal style (Update_Pos.f90:912-921).
  subroutine ibx_check(ctx, rc, where)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (rc == 0) return
    text = ''
    if (c_associated(ctx)) then
      call c_f_pointer(ibx_last_error(ctx), msg, [512])
      do i = 1, 512
        if (msg(i) == c_null_char) exit
        text(i:i) = msg(i)
      enddo
    endif
    write(6, '(a,a,a,i0,a,a)') 'libibx error in ', where, ': code ', rc, ': ', trim(text)
    error stop 'libibx failure'
  end subroutine ibx_check

end module ibx_cabi
