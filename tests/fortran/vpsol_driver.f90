! Fortran caller of the pressure-solve drop-ins, written the way the reference
! calls them: VF_VPSOL with its sizes and parameters in COMMON /VF_ANUMBI/,
! /VF_ACOMPR/ and /VF_ACOMPI/ (Source code/vf_v1cal.f:278-281), or VF_M1BCGS
! with its full argument list (vf_vpsol.f:64-69).  Both names resolve to
! libcadmas_hip.so's vf_vpsol_ / vf_m1bcgs_; the COMMON blocks declared here
! are the ones the library reads and writes.
!
!   vpsol_driver <dir> vpsol|m1bcgs
!
! <dir>/sys.txt holds NI NJ NK MI MJ MK CGPARA CGEPSA CGEPSR CGDIV ICGTYP
! ICGMAX; <dir>/{AD,ALI,ALJ,ALK,AUI,AUJ,AUK,BB}.bin the matrix as raw REAL*8
! (MI,MJ,MK) arrays and <dir>/INDC.bin the INTEGER*4 flags.  The solution goes
! to <dir>/PT.bin; ITR, !B! and !R! are printed.  VF_A2ERR below stands in for
! the reference's error exit: it prints its message and stops with code 3.
program vpsol_driver
  implicit none
  integer numi, numj, numk, numb, nums, leqk, leqt, leqc, numij, numbx
  common /vf_anumbi/ numi, numj, numk, numb, nums, leqk, leqt, leqc, numij, numbx
  integer icgtyp, icgmax, icgitr, iscmvp, iscmff, iscmk, iscmt, iscmc(10), ibsuw0
  common /vf_acompi/ icgtyp, icgmax, icgitr, iscmvp, iscmff, iscmk, iscmt, iscmc, ibsuw0
  double precision cgpara, cgepsa, cgepsr, cgdiv, cgbnrm, cgxnrm, feps, flower, fupper, fsum, &
                   fcut, plower, scmvp, scmk, scmt, scmc(10)
  common /vf_acompr/ cgpara, cgepsa, cgepsr, cgdiv, cgbnrm, cgxnrm, feps, flower, fupper, fsum, &
                     fcut, plower, scmvp, scmk, scmt, scmc
  character(len=512) :: dir, mode
  integer :: ni, nj, nk, mi, mj, mk, u, iter, ityp, imax
  double precision :: cpar, cea, cer, cdv, bn, xn
  double precision, allocatable :: a(:, :, :, :), x(:, :, :), wk(:, :, :, :), dbuf(:)
  integer, allocatable :: indc(:, :, :)
  character(len=3), parameter :: names(8) = ['AD ', 'ALI', 'ALJ', 'ALK', 'AUI', 'AUJ', 'AUK', 'BB ']
  integer :: q

  call get_command_argument(1, dir)
  call get_command_argument(2, mode)
  open (newunit=u, file=trim(dir)//'/sys.txt', status='old')
  read (u, *) ni, nj, nk, mi, mj, mk, cpar, cea, cer, cdv, ityp, imax
  close (u)
  allocate (a(mi, mj, mk, 8), x(mi, mj, mk), wk(mi, mj, mk, 8), indc(mi, mj, mk), dbuf(1))
  do q = 1, 8
    open (newunit=u, file=trim(dir)//'/'//trim(names(q))//'.bin', access='stream', status='old')
    read (u) a(:, :, :, q)
    close (u)
  end do
  open (newunit=u, file=trim(dir)//'/INDC.bin', access='stream', status='old')
  read (u) indc
  close (u)
  x = 0.0d0
  wk = 0.0d0
  if (trim(mode) == 'vpsol') then
    numi = ni
    numj = nj
    numk = nk
    cgpara = cpar
    cgepsa = cea
    cgepsr = cer
    cgdiv = cdv
    icgtyp = ityp
    icgmax = imax
    icgitr = -99
    x = -1.0d0                      ! VF_VPSOL sets PT = 0 itself
    call vf_vpsol(dbuf, a(:, :, :, 1), a(:, :, :, 2), a(:, :, :, 3), a(:, :, :, 4), &
                  a(:, :, :, 5), a(:, :, :, 6), a(:, :, :, 7), a(:, :, :, 8), x, &
                  wk(:, :, :, 1), wk(:, :, :, 2), wk(:, :, :, 3), wk(:, :, :, 4), &
                  wk(:, :, :, 5), wk(:, :, :, 6), wk(:, :, :, 7), wk(:, :, :, 8), indc)
    iter = icgitr
    bn = cgbnrm
    xn = cgxnrm
  else
    call vf_m1bcgs(dbuf, a(:, :, :, 1), a(:, :, :, 2), a(:, :, :, 3), a(:, :, :, 4), &
                   a(:, :, :, 5), a(:, :, :, 6), a(:, :, :, 7), a(:, :, :, 8), x, &
                   wk(:, :, :, 1), wk(:, :, :, 2), wk(:, :, :, 3), wk(:, :, :, 4), &
                   wk(:, :, :, 5), wk(:, :, :, 6), wk(:, :, :, 7), wk(:, :, :, 8), &
                   cpar, cea, cer, cdv, bn, xn, indc, mi, mj, mk, ni, nj, nk, ityp, imax, iter)
  end if
  open (newunit=u, file=trim(dir)//'/PT.bin', access='stream', status='replace')
  write (u) x
  close (u)
  write (*, '(a, i0, 2(1x, es24.16))') 'ITR= ', iter, bn, xn
end program vpsol_driver

! The reference's error exit (vf_a2err.f:28-49) for this driver.
subroutine vf_a2err(aname, amsg)
  implicit none
  character(len=*), intent(in) :: aname, amsg
  write (*, '(a, 1x, a, 1x, a)') 'VF_A2ERR:', trim(aname), trim(amsg)
  stop 3
end subroutine vf_a2err
