! interp_spread_ibx.f90 -- drop-in replacement for the reference's interp_spread.f90 on one
! rank, running the transfer on an MI355X through libibx.
!
! Same module name, same three public argument-less subroutines, same module state as
! interp_spread.f90:1-9: reads ap(1:pmax) (common.f90:67-88), unew/vnew/wnew, and writes
! ap(:)%ul/vl/wl, forcex/forcey/forcez, gx/gy/gz and dVeul (common.f90:6-9,84-86).  The
! caller (main.f90:225,233; indicator.f90:34) is unchanged.
!
! Results are bit-identical to the reference's single-rank loop nest: every capsule is a
! master whose eight image copies live on the same rank (interp_spread.f90:60-146), which the
! library treats as periodic images in x and y.  Capsules with mslv <= 0 (slave copies of a
! multi-rank run) get ul = 0, like interp_spread.f90:443-448, and contribute nothing to the
! spread.  Multi-rank decompositions (dims /= 1x1) need the ghost exchange of the reference
! and are rejected.
module mod_interp_spread
  use iso_c_binding
  use mod_param
  use mod_common
  use mod_common_mpi
  use ibx_cabi
  implicit none
  private
  public eulr2lagr, lagr2eulr, lagr2eulrind

  type(c_ptr), save :: ctx = c_null_ptr
  ! structure-of-arrays staging of the master capsules' points
  real(c_double), allocatable, save :: sx(:), sy(:), sz(:), s1(:), s2(:), s3(:), sa(:)
  integer(c_int64_t), allocatable, save :: soff(:)
  integer, allocatable, save :: owned(:)
  integer, save :: nown = 0

contains

  subroutine setup
    type(ibx_grid) :: g
    integer(c_int) :: rc
    if (c_associated(ctx)) return
    if (dims(1) * dims(2) /= 1) then
      write(6, *) 'mod_interp_spread (libibx): one rank only, dims = ', dims
      error stop 'libibx drop-in needs dims = 1x1'
    endif
    g%itot = itot
    g%jtot = jtot
    g%ktot = ktot
    g%reserved = 0
    g%lx = lx
    g%ly = ly
    g%lz = lz
    g%dxi = dxi
    g%dyi = dyi
    g%dzi = dzi
    g%dx = dx
    g%dy = dy
    g%dz = dz
    g%boundleft = boundleftmyid
    g%boundfront = boundfrontmyid
    rc = ibx_create(ctx, g, 0_c_int)
    call ibx_check(c_null_ptr, rc, 'ibx_create')
    allocate(sx(npmax * nl), sy(npmax * nl), sz(npmax * nl), s1(npmax * nl), &
             s2(npmax * nl), s3(npmax * nl), sa(npmax * nl), soff(npmax + 1), owned(npmax))
  end subroutine setup

  ! ap(p)%xfp(1:nl) of the masters -> contiguous arrays, capsule after capsule
  subroutine pack_positions
    integer :: p, q, n
    nown = 0
    soff(1) = 0
    do p = 1, pmax
      if (ap(p)%mslv > 0) then
        nown = nown + 1
        owned(nown) = p
        n = (nown - 1) * nl
        do q = 1, nl
          sx(n + q) = ap(p)%xfp(q)
          sy(n + q) = ap(p)%yfp(q)
          sz(n + q) = ap(p)%zfp(q)
        enddo
        soff(nown + 1) = nown * nl
      endif
    enddo
  end subroutine pack_positions

  ! eulr2lagr, interp_spread.f90:12-475
  subroutine eulr2lagr
    integer :: p, q, k, n
    integer(c_int) :: rc
    call setup
    call pack_positions
    rc = ibx_eulr2lagr(ctx, unew, vnew, wnew, int(nown, c_int64_t), soff, sx, sy, sz, s1, s2, s3)
    call ibx_check(ctx, rc, 'eulr2lagr')
    do p = 1, pmax
      ap(p)%ul(:) = 0.
      ap(p)%vl(:) = 0.
      ap(p)%wl(:) = 0.
    enddo
    do k = 1, nown
      p = owned(k)
      n = (k - 1) * nl
      do q = 1, nl
        ap(p)%ul(q) = s1(n + q)
        ap(p)%vl(q) = s2(n + q)
        ap(p)%wl(q) = s3(n + q)
      enddo
    enddo
  end subroutine eulr2lagr

  ! lagr2eulr, interp_spread.f90:510-922
  subroutine lagr2eulr
    integer :: p, q, k, n
    integer(c_int) :: rc
    call setup
    dVeul = dx * dy * dz
    call pack_positions
    do k = 1, nown
      p = owned(k)
      n = (k - 1) * nl
      do q = 1, nl
        s1(n + q) = ap(p)%fxl(q)
        s2(n + q) = ap(p)%fyl(q)
        s3(n + q) = ap(p)%fzl(q)
        sa(n + q) = ap(p)%area(q)
      enddo
    enddo
    rc = ibx_lagr2eulr(ctx, int(nown, c_int64_t), soff, sx, sy, sz, s1, s2, s3, sa, &
                       forcex, forcey, forcez)
    call ibx_check(ctx, rc, 'lagr2eulr')
  end subroutine lagr2eulr

  ! lagr2eulrind, interp_spread.f90:978-1387
  subroutine lagr2eulrind
    integer :: p, q, k, n
    integer(c_int) :: rc
    call setup
    dVeul = dx * dy * dz
    call pack_positions
    do k = 1, nown
      p = owned(k)
      n = (k - 1) * nl
      do q = 1, nl
        s1(n + q) = ap(p)%xn(q)
        s2(n + q) = ap(p)%yn(q)
        s3(n + q) = ap(p)%zn(q)
        sa(n + q) = ap(p)%area(q)
      enddo
    enddo
    rc = ibx_lagr2eulrind(ctx, int(nown, c_int64_t), soff, sx, sy, sz, s1, s2, s3, sa, gx, gy, gz)
    call ibx_check(ctx, rc, 'lagr2eulrind')
  end subroutine lagr2eulrind

end module mod_interp_spread
