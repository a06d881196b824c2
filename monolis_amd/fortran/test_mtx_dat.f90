!> Fortran caller of the GPU path: the 10x10 tridiagonal system of monolis's
!> src_test/input.f/mtx.dat (diagonal 4, off-diagonal 1), b = A*1, solved by CG and
!> BiCGSTAB with NONE and DIAG preconditioning; x = 1 expected
!> (src_test/input.f/main.f90:199-231).  Prints one line per case and
!> "FORTRAN_OK" when every case passes; stops with code 1 otherwise.
program test_mtx_dat
  use iso_c_binding
  use mod_monolis_hip
  implicit none
  integer(c_int), parameter :: N = 10
  integer(c_int) :: index(N + 1), item(3*N), Iarray(100)
  real(c_double) :: A(3*N), x(N), b(N), one(N), Rarray(100), s
  integer(c_int) :: i, k, nz, method, prec
  logical :: ok

  ! CSR of the symmetric tridiagonal matrix, diagonal included, ascending columns
  nz = 0
  index(1) = 0
  do i = 1, N
    if (i > 1) then
      nz = nz + 1; item(nz) = i - 1; A(nz) = 1.0d0
    endif
    nz = nz + 1; item(nz) = i; A(nz) = 4.0d0
    if (i < N) then
      nz = nz + 1; item(nz) = i + 1; A(nz) = 1.0d0
    endif
    index(i + 1) = nz
  enddo

  one = 1.0d0
  call monolis_hip_matvec_R(N, 1, index, item, A, one, b)
  call monolis_hip_inner_product_R(N, one, b, s)
  ok = abs(s - 58.0d0) < 1.0d-12   ! sum of row sums: 2*5 + 8*6
  write(*,"(a,f8.3)") "inner product (1, A 1) = ", s

  do k = 0, 3
    method = merge(monolis_hip_iter_CG, monolis_hip_iter_BiCGSTAB, k < 2)
    prec = merge(monolis_hip_prec_NONE, monolis_hip_prec_DIAG, mod(k, 2) == 0)
    call monolis_hip_prm_initialize(Iarray, Rarray)
    Iarray(1) = method
    Iarray(2) = prec
    Iarray(20) = 0
    Iarray(21) = 0
    Iarray(22) = 0
    Rarray(1) = 1.0d-10
    x = 0.0d0
    call monolis_hip_solve_R(N, 1, index, item, A, b, x, Iarray, Rarray)
    write(*,"(a,i2,a,i2,a,i4,a,1pe10.3,a,1pe10.3)") "method", method, " prec", prec, &
      " iter", Iarray(4), " resid", Rarray(2), " max|x-1|", maxval(abs(x - 1.0d0))
    if (Iarray(5) /= 0 .or. Iarray(4) < 1 .or. Rarray(2) > 1.0d-10) ok = .false.
    if (maxval(abs(x - 1.0d0)) > 1.0d-9) ok = .false.
  enddo

  if (ok) then
    write(*,"(a)") "FORTRAN_OK"
  else
    stop 1
  endif
end program test_mtx_dat
