! Drop-in caller of the legacy symbol DPJDREVCOM (reverse communication,
! diagonal preconditioner selected by JA(1) < 0) and DPJDCLEANUP, compiled
! with ROCm flang and linked against libjadamilu_amd.so exactly as a program
! using the reference library would be.  Matrix: tridiag(5, i, 5), N = 1000,
! 5 smallest eigenpairs, TOL 1e-10 (the setting of the reference's diagonal
! preconditioning sample).  Prints the eigenvalues, residuals and MATVEC count.
program dropin_revcom
  implicit none
  integer, parameter :: n = 1000, neigmax = 5, msp = 20
  integer, parameter :: lx = n*(3*msp + neigmax + 1) + 4*msp*msp
  double precision :: eigs(neigmax), res(neigmax), x(lx), d(n)
  double precision :: sigma, tol, gap, mem, droptol, shift
  integer :: neig, madspace, isearch, ninit, icntl(5), iter, iprint, info
  integer :: ijob, ndx1, ndx2, i, ja1(1), ia1(1)

  do i = 1, n
     d(i) = dble(i)
  end do
  x = 0.0d0
  eigs = 0.0d0
  res = 0.0d0
  sigma = 0.0d0
  shift = 0.0d0
  droptol = 0.0d0
  mem = 0.0d0
  iprint = 0
  isearch = 0
  neig = neigmax
  ninit = 0
  madspace = msp
  iter = 1000
  tol = 1.0d-10
  icntl = 0
  ja1(1) = -1
  ia1(1) = -1
  ijob = 0
  do
     call dpjdrevcom(n, d, ja1, ia1, eigs, res, x, lx, neig, sigma, isearch, ninit, &
                     madspace, iter, tol, shift, droptol, mem, icntl, ijob, ndx1, ndx2, &
                     iprint, info, gap)
     if (ijob /= 1) exit
     call tridiag_mv(n, x(ndx1:ndx1+n-1), x(ndx2:ndx2+n-1))
  end do
  call dpjdcleanup()
  print '(A,I6,A,I7)', 'MATVEC', iter, ' INFO', info
  do i = 1, neig
     print '(ES24.16,ES12.3)', eigs(i), res(i)
  end do
contains
  subroutine tridiag_mv(m, xin, yout)
    integer, intent(in) :: m
    double precision, intent(in) :: xin(m)
    double precision, intent(out) :: yout(m)
    integer :: k
    do k = 1, m
       yout(k) = dble(k)*xin(k)
       if (k > 1) yout(k) = yout(k) + 5.0d0*xin(k-1)
       if (k < m) yout(k) = yout(k) + 5.0d0*xin(k+1)
    end do
  end subroutine tridiag_mv
end program dropin_revcom
