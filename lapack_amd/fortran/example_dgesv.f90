! example_dgesv.f90 -- a Fortran caller of the library, two ways:
!   1. unchanged reference-style calls (CALL DGETRF / CALL DGETRS / CALL DSGESV),
!      resolved at link time to liblapack_amd's Fortran symbols;
!   2. the ISO_C_BINDING module (INFO as a function result, DGESV_IR in BF16).
! Prints one line per check and "OK" at the end when every residual ratio
! ||b - A x||_inf / (||A||_inf ||x||_inf eps) is below 30 (the reference test
! suite's threshold).
program example_dgesv
  use lapack_amd_mod
  implicit none
  integer, parameter :: n = 512, nrhs = 2
  double precision :: a(n, n), a0(n, n), b(n, nrhs), x(n, nrhs), berr(nrhs)
  integer :: ipiv(n), info, i, j, iter, giter
  double precision :: r1, r2, r3, r4
  logical :: ok
  external :: dgetrf, dgetrs, dsgesv
  double precision :: work(n*nrhs)
  real :: swork(n*(n+nrhs))

  do j = 1, n
    do i = 1, n
      a0(i, j) = sin(dble(i * 7 + j * 13)) + merge(dble(n) / 8.0d0, 0.0d0, i == j)
    end do
  end do
  do j = 1, nrhs
    do i = 1, n
      b(i, j) = cos(dble(i + 3 * j))
    end do
  end do

  ! 1. reference calling convention
  a = a0
  x = b
  call dgetrf(n, n, a, n, ipiv, info)
  if (info /= 0) error stop 'dgetrf failed'
  call dgetrs('N', n, nrhs, a, n, ipiv, x, n, info)
  if (info /= 0) error stop 'dgetrs failed'
  r1 = ratio(a0, x, b)
  print '(a, es10.3)', 'DGETRF + DGETRS (drop-in symbols):  ratio ', r1

  a = a0
  call dsgesv(n, nrhs, a, n, ipiv, b, n, x, n, work, swork, iter, info)
  if (info /= 0) error stop 'dsgesv failed'
  r2 = ratio(a0, x, b)
  print '(a, es10.3, a, i3)', 'DSGESV (drop-in symbol):            ratio ', r2, '  ITER ', iter

  ! 2. ISO_C_BINDING module
  a = a0
  x = b
  info = lapack_amd_dgesv(n, nrhs, a, n, ipiv, x, n)
  if (info /= 0) error stop 'lapack_amd_dgesv failed'
  r3 = ratio(a0, x, b)
  print '(a, es10.3)', 'lapack_amd_dgesv (bind(C)):         ratio ', r3

  a = a0
  info = lapack_amd_dgesv_ir(n, nrhs, a, n, ipiv, b, n, x, n, 'B', iter, giter, berr)
  if (info /= 0) error stop 'lapack_amd_dgesv_ir failed'
  r4 = ratio(a0, x, b)
  print '(a, es10.3, a, i3, a, i3)', 'lapack_amd_dgesv_ir BF16 (bind(C)): ratio ', r4, '  ITER ', iter, &
      '  GMRES ', giter

  ok = max(r1, r2, r3, r4) < 30.0d0 .and. iter >= 0
  if (ok) then
    print '(a)', 'OK'
  else
    print '(a)', 'FAILED'
    error stop 1
  end if

contains

  double precision function ratio(am, xm, bm)
    double precision, intent(in) :: am(n, n), xm(n, nrhs), bm(n, nrhs)
    double precision :: r(n), anorm, xnorm, rnorm
    integer :: k
    anorm = maxval(sum(abs(am), dim=2))
    ratio = 0.0d0
    do k = 1, nrhs
      r = bm(:, k) - matmul(am, xm(:, k))
      rnorm = maxval(abs(r))
      xnorm = maxval(abs(xm(:, k)))
      ratio = max(ratio, rnorm / (anorm * xnorm * epsilon(1.0d0) * 0.5d0))
    end do
  end function

end program example_dgesv
