! lapack_amd_mod.f90 -- ISO_C_BINDING interface to the value-argument entry
! points of liblapack_amd (include/lapack_amd.h).  Unchanged Fortran code needs
! none of this: it calls DGETRF/DGETRS/DGESV/DSGESV/DLASWP by name and links
! against liblapack_amd, which exports those symbols with the reference's
! calling convention.  This module is for code that wants INFO as a function
! result, the device (GPU-resident) arrays, or the mixed-precision solver.
module lapack_amd_mod
  use iso_c_binding
  implicit none
  interface
    integer(c_int) function lapack_amd_dgetrf(m, n, a, lda, ipiv) bind(C, name="lapack_amd_dgetrf")
      import :: c_int, c_double
      integer(c_int), value :: m, n, lda
      real(c_double) :: a(lda, *)
      integer(c_int) :: ipiv(*)
    end function
    integer(c_int) function lapack_amd_dgetrs(trans, n, nrhs, a, lda, ipiv, b, ldb) &
        bind(C, name="lapack_amd_dgetrs")
      import :: c_int, c_double, c_char
      character(kind=c_char), value :: trans
      integer(c_int), value :: n, nrhs, lda, ldb
      real(c_double), intent(in) :: a(lda, *)
      integer(c_int), intent(in) :: ipiv(*)
      real(c_double) :: b(ldb, *)
    end function
    integer(c_int) function lapack_amd_dgesv(n, nrhs, a, lda, ipiv, b, ldb) bind(C, name="lapack_amd_dgesv")
      import :: c_int, c_double
      integer(c_int), value :: n, nrhs, lda, ldb
      real(c_double) :: a(lda, *), b(ldb, *)
      integer(c_int) :: ipiv(*)
    end function
    integer(c_int) function lapack_amd_dsgesv(n, nrhs, a, lda, ipiv, b, ldb, x, ldx, iter) &
        bind(C, name="lapack_amd_dsgesv")
      import :: c_int, c_double
      integer(c_int), value :: n, nrhs, lda, ldb, ldx
      real(c_double) :: a(lda, *), x(ldx, *)
      real(c_double), intent(in) :: b(ldb, *)
      integer(c_int) :: ipiv(*), iter
    end function
    integer(c_int) function lapack_amd_dgesv_ir(n, nrhs, a, lda, ipiv, b, ldb, x, ldx, prec, iter, giter, berr) &
        bind(C, name="lapack_amd_dgesv_ir")
      import :: c_int, c_double, c_char
      integer(c_int), value :: n, nrhs, lda, ldb, ldx
      real(c_double) :: a(lda, *), x(ldx, *), berr(*)
      real(c_double), intent(in) :: b(ldb, *)
      integer(c_int) :: ipiv(*), iter, giter
      character(kind=c_char), value :: prec
    end function
  end interface
end module lapack_amd_mod
