! lin_ge_main.f90 -- drives the reference's own test routines of the GE path
! (TESTING/LIN: DCHKGE with DERRGE, DDRVGE with DERRVX, DDRVAB with DERRAB)
! against whatever DGETRF / DGETRS / DGESV / DSGESV / DGERFS / DGEEQU /
! DLAQGE the program is linked with -- here the GPU library's drop-in
! symbols (liblapack_amd.so), the rest of LAPACK and the test routines coming
! from the reference sources.  Test infrastructure: built by oracle/Makefile
! into oracle/_ref/ and run by tests/test_gpu_reference_suite.py.
!
! Parameter sets: those of TESTING/dtest.in:2-16 for the GE path (M and N in
! {0,1,2,3,5,10,50}, NRHS {1,2,15}, NB {1,3,20}, threshold 30, error exits
! on, all 11 matrix types) and of TESTING/dstest.in:2-8 for DSGESV (M in
! {0,1,2,13,17,45,78,91,101,120,132}, NRHS {1,2,15,16}, threshold 30).
! The test routines report through ALASUM / ALASVM / ALAERH on unit 6.
program lin_ge
  implicit none
  integer, parameter :: nmax = 132, maxin = 12, maxrhs = 16, ntypes = 11
  integer, parameter :: kdmax = nmax + (nmax + 1) / 4, nout = 6
  ! the test routines' shared state (TESTING/LIN/xerbla.f, chkxer.f, xlaenv.f)
  integer :: infot, nunit
  logical :: ok, lerr
  character(len=32) :: srnamt
  integer :: iparms(100)
  common /infoc/ infot, nunit, ok, lerr
  common /srnamc/ srnamt
  common /claenv/ iparms

  logical :: dotype(30)
  integer :: mval(maxin), nval(maxin), nsval(maxin), nbval(maxin)
  integer :: mval2(maxin), nsval2(maxin)
  integer :: nm, nn, nns, nnb, nm2, nns2
  double precision :: thresh
  double precision, allocatable :: a(:, :), b(:, :), work(:, :), rwork(:), s(:)
  double precision, allocatable :: a2(:, :), b2(:, :), work2(:), rwork2(:)
  real, allocatable :: swork(:)
  integer, allocatable :: iwork(:), iwork2(:)

  iparms = 0
  nunit = nout
  ok = .true.
  lerr = .false.
  infot = 0
  srnamt = ' '
  dotype = .true.
  thresh = 30.0d0

  nm = 7
  mval(1:7) = (/ 0, 1, 2, 3, 5, 10, 50 /)
  nn = 7
  nval(1:7) = (/ 0, 1, 2, 3, 5, 10, 50 /)
  nns = 3
  nsval(1:3) = (/ 1, 2, 15 /)
  nnb = 3
  nbval(1:3) = (/ 1, 3, 20 /)

  allocate (a((kdmax + 1) * nmax, 8), b(nmax * maxrhs, 4))
  allocate (work(nmax, 3 * nmax + maxrhs + 30), rwork(5 * nmax + 2 * maxrhs), s(2 * nmax))
  allocate (iwork(50 * nmax))

  write (nout, '(a)') ' GE path: DCHKGE (DGETRF, DGETRS, DGERFS, DGETRI, DGECON) and DDRVGE (DGESV, DGESVX)'
  call dchkge(dotype, nm, mval, nn, nval, nnb, nbval, nns, nsval, thresh, .true., nmax, &
              a(1, 1), a(1, 2), a(1, 3), b(1, 1), b(1, 2), b(1, 3), work, rwork, iwork, nout)
  call ddrvge(dotype, nn, nval, nsval(1), thresh, .true., nmax, a(1, 1), a(1, 2), a(1, 3), &
              b(1, 1), b(1, 2), b(1, 3), b(1, 4), s, work, rwork, iwork, nout)

  nm2 = 11
  mval2(1:11) = (/ 0, 1, 2, 13, 17, 45, 78, 91, 101, 120, 132 /)
  nns2 = 4
  nsval2(1:4) = (/ 1, 2, 15, 16 /)
  allocate (a2(nmax * nmax, 2), b2(nmax * maxrhs, 2), work2(nmax * maxrhs * 2), rwork2(nmax))
  allocate (swork(nmax * (nmax + maxrhs)), iwork2(nmax))
  write (nout, '(a)') ' DSGESV: DERRAB and DDRVAB'
  call derrab(nout)
  call ddrvab(dotype, nm2, mval2, nns2, nsval2, thresh, nmax, a2(1, 1), a2(1, 2), b2(1, 1), b2(1, 2), &
              work2, rwork2, swork, iwork2, nout)
  write (nout, '(a)') ' END OF GE TESTS'
end program lin_ge
