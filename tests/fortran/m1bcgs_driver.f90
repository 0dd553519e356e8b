! Fortran caller of the VF_M1BCGS drop-in (cadmas_hip_m1bcgs_), as the
! reference's VF_VPSOL would call it (Source code/vf_vpsol.f:64-67): a 7-point
! system with the reference's sign convention (AD > 0, off-diagonals <= 0 on
! active rows, the AL*/AU* of a neighbour that is inactive equal to 0), solved
! to the reference's stop test; the driver checks the true residual in FP64.
program m1bcgs_driver
  use cadmas_hip
  implicit none
  integer, parameter :: ni = 20, nj = 12, nk = 10
  real(8), dimension(ni, nj, nk) :: ad, ali, alj, alk, aui, auj, auk, bb, x, r
  integer :: indc(ni, nj, nk), iter, i, j, k
  real(8) :: bn, xn, rn, b2, s
  indc = -1
  indc(2:ni-1, 2:nj-1, 2:nk-1) = 0
  indc(5:8, 4:6, 2:4) = -1                       ! an obstacle block
  ad = 1; ali = 0; alj = 0; alk = 0; aui = 0; auj = 0; auk = 0; bb = 0; x = 0
  do k = 2, nk - 1
    do j = 2, nj - 1
      do i = 2, ni - 1
        if (indc(i, j, k) == -1) cycle
        if (indc(i-1, j, k) /= -1) ali(i, j, k) = -1.0d0 - 0.01d0 * i
        if (indc(i+1, j, k) /= -1) aui(i, j, k) = -1.0d0 - 0.01d0 * (i + 1)
        if (indc(i, j-1, k) /= -1) alj(i, j, k) = -0.5d0
        if (indc(i, j+1, k) /= -1) auj(i, j, k) = -0.5d0
        if (indc(i, j, k-1) /= -1) alk(i, j, k) = -2.0d0
        if (indc(i, j, k+1) /= -1) auk(i, j, k) = -2.0d0
        ad(i, j, k) = 0.1d0 - (ali(i,j,k) + aui(i,j,k) + alj(i,j,k) + auj(i,j,k) + alk(i,j,k) + auk(i,j,k))
        bb(i, j, k) = sin(0.3d0 * i) + cos(0.2d0 * j * k)
      end do
    end do
  end do
  call cadmas_hip_m1bcgs(ad, ali, alj, alk, aui, auj, auk, bb, x, 0.95d0, 1.0d-15, 1.0d-12, 1.0d30, &
                         bn, xn, indc, ni, nj, nk, ni, nj, nk, 1, 500, iter)
  rn = 0; b2 = 0
  do k = 2, nk - 1
    do j = 2, nj - 1
      do i = 2, ni - 1
        if (indc(i, j, k) == -1) cycle
        s = ad(i,j,k)*x(i,j,k) + ali(i,j,k)*x(i-1,j,k) + aui(i,j,k)*x(i+1,j,k) &
          + alj(i,j,k)*x(i,j-1,k) + auj(i,j,k)*x(i,j+1,k) + alk(i,j,k)*x(i,j,k-1) + auk(i,j,k)*x(i,j,k+1)
        r(i, j, k) = bb(i, j, k) - s
        rn = rn + r(i, j, k)**2
        b2 = b2 + bb(i, j, k)**2
      end do
    end do
  end do
  write (*, '(a, i4, a, es12.4)') 'M1BCGS ITER=', iter, '  true relative residual=', sqrt(rn / b2)
  if (iter < 1 .or. iter > 500 .or. sqrt(rn / b2) > 1.0d-8) then
    write (*, '(a)') 'FAIL'
    stop 1
  end if
  write (*, '(a)') 'OK'
end program m1bcgs_driver
