! test_dropin.f90 -- calls the drop-in eulr2lagr / lagr2eulr / lagr2eulrind exactly as the
! reference's main program does (main.f90:225,233; indicator.f90:34), on module state read
! from a stream file, and writes the module state they produce.  TEST INFRASTRUCTURE.
!   usage: test_dropin <input file> <output file>
program test_dropin
  use mod_param
  use mod_common
  use mod_interp_spread
  implicit none
  character(len=1024) :: fin, fout
  integer :: p, uin, uout
  integer(4) :: np4
  integer(4) :: ms(npmax)
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=uin, file=trim(fin), access='stream', form='unformatted', status='old')
  read(uin) np4
  pmax = np4
  read(uin) ms(1:pmax)
  read(uin) unew
  read(uin) vnew
  read(uin) wnew
  do p = 1, pmax
    ap(p)%mslv = ms(p)
    read(uin) ap(p)%xfp, ap(p)%yfp, ap(p)%zfp, ap(p)%fxl, ap(p)%fyl, ap(p)%fzl, ap(p)%area, &
              ap(p)%xn, ap(p)%yn, ap(p)%zn
  enddo
  close(uin)
  call eulr2lagr
  call lagr2eulr
  call lagr2eulrind
  open(newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace')
  do p = 1, pmax
    write(uout) ap(p)%ul, ap(p)%vl, ap(p)%wl
  enddo
  write(uout) forcex, forcey, forcez, gx, gy, gz, dVeul
  close(uout)
  write(6, '(a)') 'test_dropin: done'
end program test_dropin
