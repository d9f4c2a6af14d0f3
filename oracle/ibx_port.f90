! ibx_port.f90 -- the reference's single-rank interpolation and spread loops restated in
! Fortran, for the CPU baseline of bench.py (and checked against the C oracle by
! tests/test_cpu_port.py).  TEST AND MEASUREMENT INFRASTRUCTURE ONLY: the product never links,
! calls or falls back to it.
!
!   port_eulr2lagr   interp_spread.f90:12-475  (images :60-146, partial integration
!                    :300-383 with its OpenMP loop over capsules, owner sum :426-440)
!   port_lagr2eulr   interp_spread.f90:510-922 (zeroing :811-815, partial integration
!                    :817-884 with its OpenMP loop over capsules and atomic updates,
!                    halos :904-912)
!
! One rank, so every capsule is a master that is its own neighbour in every direction: image
! nb (1..8) is formed when some point lies within `offset` of the matching boundary
! (initparticles.f90:241-295, offset from param.f90:135) and is stored in slot nb+4 (mod 8)
! shifted by +-lx / +-ly (interp_spread.f90:82-141, 333-336).  The arithmetic is the
! reference's, term by term (kernel.f90:7-25; products left to right), so the serial build
! gives the C oracle's bits.  Built as the reference ships it (-O0, OpenMP off, Makefile:7,14)
! and at -O2, each with and without OpenMP (-fopenmp turns on the reference's own directives:
! the spread then adds with atomics, in a thread-dependent order).
module ibx_port
  use iso_c_binding
  implicit none
  private
  public :: port_eulr2lagr, port_lagr2eulr, port_max_threads, port_set_threads
  integer, parameter :: dp = c_double
  ! slot s -> shift in units of (lx, ly): interp_spread.f90:82-141
  integer, parameter :: sx(0:8) = (/ 0, 1, 1, 0, -1, -1, -1, 0, 1 /)
  integer, parameter :: sy(0:8) = (/ 0, 0, -1, -1, -1, 0, 1, 1, 1 /)

contains

  ! kernel.f90:7-25 (the delta function times h)
  pure real(dp) function kern(r)
    real(dp), intent(in) :: r
    real(dp) :: a
    a = abs(r)
    if (a > 1.5_dp) then
      kern = 0._dp
    else if (a > 0.5_dp) then
      kern = (1._dp / 6._dp) * (5._dp - 3._dp * a - sqrt(-3._dp * ((1._dp - a)**2) + 1._dp))
    else
      kern = (1._dp / 3._dp) * (1._dp + sqrt(-3._dp * (r**2) + 1._dp))
    end if
  end function kern

  ! the images the reference forms for one capsule (initparticles.f90:241-295 on one rank:
  ! left/front bounds 0, right/back bounds lx/ly)
  pure subroutine images(n, x, y, lx, ly, off, nbf)
    integer(c_int64_t), intent(in) :: n
    real(dp), intent(in) :: x(n), y(n), lx, ly, off
    logical, intent(out) :: nbf(8)
    integer(c_int64_t) :: l
    nbf = .false.
    do l = 1, n
      if (x(l) + off > lx) then
        if (y(l) + off >= 0._dp .and. y(l) - off <= ly) nbf(1) = .true.
        if (y(l) - off <= 0._dp) nbf(2) = .true.
      end if
      if (y(l) - off < 0._dp) then
        if (x(l) + off >= 0._dp .and. x(l) - off <= lx) nbf(3) = .true.
        if (x(l) - off <= 0._dp) nbf(4) = .true.
      end if
      if (x(l) - off < 0._dp) then
        if (y(l) + off >= 0._dp .and. y(l) - off <= ly) nbf(5) = .true.
      end if
      if (y(l) + off > ly) then
        if (x(l) - off <= 0._dp) nbf(6) = .true.
        if (x(l) + off >= 0._dp .and. x(l) - off <= lx) nbf(7) = .true.
        if (x(l) + off >= lx) nbf(8) = .true.
      end if
    end do
  end subroutine images

  pure integer function slot_of(nb)
    integer, intent(in) :: nb
    slot_of = nb
    if (nb > 0) then
      slot_of = nb + 4
      if (slot_of > 8) slot_of = slot_of - 8
    end if
  end function slot_of

  integer(c_int) function port_max_threads() bind(c, name="port_max_threads")
    !$ integer, external :: omp_get_max_threads
    port_max_threads = 1
    !$ port_max_threads = omp_get_max_threads()
  end function port_max_threads

  subroutine port_set_threads(n) bind(c, name="port_set_threads")
    integer(c_int), value :: n
    !$ call omp_set_num_threads(max(n, 1))
  end subroutine port_set_threads

  ! grid: itot, jtot, ktot; lx, ly, dxi, dyi, dzi (boundleft = boundfront = 0)
  subroutine port_eulr2lagr(itot, jtot, ktot, lx, ly, dxi, dyi, dzi, u, v, w, ncap, cap_off, &
                            x, y, z, ul, vl, wl) bind(c, name="port_eulr2lagr")
    integer(c_int), value :: itot, jtot, ktot
    real(dp), value :: lx, ly, dxi, dyi, dzi
    real(dp), intent(in) :: u(0:itot+1, 0:jtot+1, 0:ktot+1), v(0:itot+1, 0:jtot+1, 0:ktot+1), &
                            w(0:itot+1, 0:jtot+1, 0:ktot+1)
    integer(c_int64_t), value :: ncap
    integer(c_int64_t), intent(in) :: cap_off(0:ncap)
    real(dp), intent(in) :: x(*), y(*), z(*)
    real(dp), intent(inout) :: ul(*), vl(*), wl(*)
    integer(c_int64_t) :: p, l, a, n
    integer :: nb, s, i, j, k, ilow, ihigh, jlow, jhigh, klow, khigh
    real(dp) :: xs, ys, cx, cy, cz, czs, czz, kzs, kz, cys, cyy, kys, ky, cxs, cxx, kxs, kx, off
    real(dp) :: su, sv, sw
    real(dp), allocatable :: an(:, :, :)
    logical :: nbf(8)
    off = sqrt(3._dp * (1.5_dp**2)) / dxi + 0.01_dp / dxi
    !$omp parallel do default(shared) schedule(dynamic, 4) &
    !$omp& private(p, l, a, n, nb, s, i, j, k, ilow, ihigh, jlow, jhigh, klow, khigh) &
    !$omp& private(xs, ys, cx, cy, cz, czs, czz, kzs, kz, cys, cyy, kys, ky, cxs, cxx, kxs, kx) &
    !$omp& private(su, sv, sw, an, nbf)
    do p = 1, ncap
      a = cap_off(p - 1)
      n = cap_off(p) - a
      if (n <= 0) cycle
      allocate(an(3, n, 0:8))
      an = 0._dp
      call images(n, x(a + 1:a + n), y(a + 1:a + n), lx, ly, off, nbf)
      do nb = 0, 8
        if (nb > 0) then
          if (.not. nbf(nb)) cycle
        end if
        s = slot_of(nb)
        do l = 1, n
          xs = x(a + l)
          ys = y(a + l)
          if (sx(s) > 0) xs = xs + lx
          if (sx(s) < 0) xs = xs - lx
          if (sy(s) > 0) ys = ys + ly
          if (sy(s) < 0) ys = ys - ly
          cz = z(a + l) * dzi
          cx = (xs - 0._dp) * dxi
          cy = (ys - 0._dp) * dyi
          ilow = max(nint(cx - 2._dp), 1)
          ihigh = min(nint(cx + 2._dp), itot)
          jlow = max(nint(cy - 2._dp), 1)
          jhigh = min(nint(cy + 2._dp), jtot)
          klow = max(nint(cz - 2._dp), 1)
          khigh = min(nint(cz + 2._dp), ktot)
          do k = klow, khigh
            czs = (1._dp * k) - cz
            czz = czs - 0.5_dp
            kzs = kern(czs)
            kz = kern(czz)
            do j = jlow, jhigh
              cys = (1._dp * j) - cy
              cyy = cys - 0.5_dp
              kys = kern(cys)
              ky = kern(cyy)
              do i = ilow, ihigh
                cxs = (1._dp * i) - cx
                cxx = cxs - 0.5_dp
                kxs = kern(cxs)
                kx = kern(cxx)
                an(1, l, s) = an(1, l, s) + u(i, j, k) * kxs * ky * kz
                an(2, l, s) = an(2, l, s) + v(i, j, k) * kx * kys * kz
                an(3, l, s) = an(3, l, s) + w(i, j, k) * kx * ky * kzs
              end do
            end do
          end do
        end do
      end do
      ! the owner's sum over the nine slots, slot order (interp_spread.f90:426-440)
      do l = 1, n
        su = 0._dp
        sv = 0._dp
        sw = 0._dp
        do s = 0, 8
          su = su + an(1, l, s)
          sv = sv + an(2, l, s)
          sw = sw + an(3, l, s)
        end do
        ul(a + l) = su
        vl(a + l) = sv
        wl(a + l) = sw
      end do
      deallocate(an)
    end do
    !$omp end parallel do
  end subroutine port_eulr2lagr

  subroutine halos(itot, jtot, ktot, f)
    integer(c_int), intent(in) :: itot, jtot, ktot
    real(dp), intent(inout) :: f(0:itot+1, 0:jtot+1, 0:ktot+1)
    f(itot + 1, :, :) = f(1, :, :)
    f(0, :, :) = f(itot, :, :)
  end subroutine halos

  subroutine halos_y(itot, jtot, ktot, f)
    integer(c_int), intent(in) :: itot, jtot, ktot
    real(dp), intent(inout) :: f(0:itot+1, 0:jtot+1, 0:ktot+1)
    f(:, jtot + 1, :) = f(:, 1, :)
    f(:, 0, :) = f(:, jtot, :)
  end subroutine halos_y

  subroutine port_lagr2eulr(itot, jtot, ktot, lx, ly, dxi, dyi, dzi, dveul, ncap, cap_off, &
                            x, y, z, fx, fy, fz, area, forcex, forcey, forcez) &
                            bind(c, name="port_lagr2eulr")
    integer(c_int), value :: itot, jtot, ktot
    real(dp), value :: lx, ly, dxi, dyi, dzi, dveul
    integer(c_int64_t), value :: ncap
    integer(c_int64_t), intent(in) :: cap_off(0:ncap)
    real(dp), intent(in) :: x(*), y(*), z(*), fx(*), fy(*), fz(*), area(*)
    real(dp), intent(inout) :: forcex(0:itot+1, 0:jtot+1, 0:ktot+1), &
                               forcey(0:itot+1, 0:jtot+1, 0:ktot+1), &
                               forcez(0:itot+1, 0:jtot+1, 0:ktot+1)
    integer(c_int64_t) :: p, l, a, n, q
    integer :: nb, s, i, j, k, ilow, ihigh, jlow, jhigh, klow, khigh
    real(dp) :: xs, ys, cx, cy, cz, czs, czz, kzs, kz, cys, cyy, kys, ky, cxs, cxx, kxs, kx, off
    logical :: nbf(8)
    off = sqrt(3._dp * (1.5_dp**2)) / dxi + 0.01_dp / dxi
    !$omp parallel do default(shared) private(k)
    do k = 0, ktot + 1
      forcex(:, :, k) = 0._dp
      forcey(:, :, k) = 0._dp
      forcez(:, :, k) = 0._dp
    end do
    !$omp end parallel do
    !$omp parallel do default(shared) schedule(dynamic, 4) &
    !$omp& private(p, l, a, n, q, nb, s, i, j, k, ilow, ihigh, jlow, jhigh, klow, khigh) &
    !$omp& private(xs, ys, cx, cy, cz, czs, czz, kzs, kz, cys, cyy, kys, ky, cxs, cxx, kxs, kx) &
    !$omp& private(nbf)
    do p = 1, ncap
      a = cap_off(p - 1)
      n = cap_off(p) - a
      if (n <= 0) cycle
      call images(n, x(a + 1:a + n), y(a + 1:a + n), lx, ly, off, nbf)
      do nb = 0, 8
        if (nb > 0) then
          if (.not. nbf(nb)) cycle
        end if
        s = slot_of(nb)
        do l = 1, n
          q = a + l
          xs = x(q)
          ys = y(q)
          if (sx(s) > 0) xs = xs + lx
          if (sx(s) < 0) xs = xs - lx
          if (sy(s) > 0) ys = ys + ly
          if (sy(s) < 0) ys = ys - ly
          cx = (xs - 0._dp) * dxi
          cy = (ys - 0._dp) * dyi
          cz = z(q) * dzi
          ilow = max(nint(cx - 2._dp), 1)
          ihigh = min(nint(cx + 2._dp), itot)
          jlow = max(nint(cy - 2._dp), 1)
          jhigh = min(nint(cy + 2._dp), jtot)
          klow = max(nint(cz - 2._dp), 1)
          khigh = min(nint(cz + 2._dp), ktot)
          do k = klow, khigh
            czs = (1._dp * k) - cz
            czz = czs - 0.5_dp
            kzs = kern(czs)
            kz = kern(czz)
            do j = jlow, jhigh
              cys = (1._dp * j) - cy
              cyy = cys - 0.5_dp
              kys = kern(cys)
              ky = kern(cyy)
              do i = ilow, ihigh
                cxs = (1._dp * i) - cx
                cxx = cxs - 0.5_dp
                kxs = kern(cxs)
                kx = kern(cxx)
                !$omp atomic
                forcex(i, j, k) = forcex(i, j, k) + fx(q) * kxs * ky * kz * (area(q) / dveul)
                !$omp atomic
                forcey(i, j, k) = forcey(i, j, k) + fy(q) * kx * kys * kz * (area(q) / dveul)
                !$omp atomic
                forcez(i, j, k) = forcez(i, j, k) + fz(q) * kx * ky * kzs * (area(q) / dveul)
              end do
            end do
          end do
        end do
      end do
    end do
    !$omp end parallel do
    call halos(itot, jtot, ktot, forcex)
    call halos(itot, jtot, ktot, forcey)
    call halos(itot, jtot, ktot, forcez)
    call halos_y(itot, jtot, ktot, forcex)
    call halos_y(itot, jtot, ktot, forcey)
    call halos_y(itot, jtot, ktot, forcez)
  end subroutine port_lagr2eulr

end module ibx_port
