!> Fortran caller of the complex GPU path in the shape of monolis's
!> src_test/iterative/COCG_test.f90:41-85: a 10-node chain with 2 DOF per node,
!> complex symmetric (A_ij = A_ji, no conjugation), diagonal (2 + 2i) + u + i v,
!> b = A (1 + i), COCG with NONE and DIAG preconditioning to 1e-12; x = 1 + i
!> expected.  The random values of the reference test are replaced by a fixed
!> deterministic sequence.  Prints "FORTRAN_COCG_OK" when both cases pass.
program test_cocg_chain
  use iso_c_binding
  use mod_monolis_hip
  implicit none
  integer(c_int), parameter :: N = 10, ND = 2
  integer(c_int) :: index(N + 1), item(3*N), Iarray(100)
  complex(c_double_complex) :: A(4*3*N), x(N*ND), b(N*ND), one(N*ND), s
  real(c_double) :: Rarray(100), u, v
  integer(c_int) :: i, j, k, nz, prec, kd, ko, kt, seq
  logical :: ok

  ! chain pattern: rows hold {i-1, i, i+1}, 1-based columns
  nz = 0
  index(1) = 0
  do i = 1, N
    if (i > 1) then
      nz = nz + 1; item(nz) = i - 1
    endif
    nz = nz + 1; item(nz) = i
    if (i < N) then
      nz = nz + 1; item(nz) = i + 1
    endif
    index(i + 1) = nz
  enddo
  A = (0.0d0, 0.0d0)
  seq = 0
  do i = 1, N
    kd = index(i) + merge(1, 2, i == 1)          ! diagonal block of row i
    do j = 1, ND
      u = u01(seq)
      v = u01(seq)
      A(4*(kd - 1) + ND*(j - 1) + j) = (2.0d0, 2.0d0) + cmplx(u, v, kind=c_double)
    enddo
  enddo
  do i = 1, N - 1
    ko = index(i) + merge(2, 3, i == 1)           ! block (i, i+1)
    kt = index(i + 1) + 1                         ! block (i+1, i)
    do j = 1, ND
      do k = 1, ND
        u = u01(seq)
        v = u01(seq)
        A(4*(ko - 1) + ND*(j - 1) + k) = cmplx(u, v, kind=c_double)
        A(4*(kt - 1) + ND*(k - 1) + j) = A(4*(ko - 1) + ND*(j - 1) + k)
      enddo
    enddo
  enddo

  one = (1.0d0, 1.0d0)
  call monolis_hip_matvec_C(N, ND, index, item, A, one, b)
  call monolis_hip_inner_product_C(N*ND, one, one, s)
  ok = abs(s - (0.0d0, 40.0d0)) < 1.0d-12        ! (1+i)^2 = 2i per entry
  write(*,"(a,2f9.3)") "inner product (1+i, 1+i) = ", real(s), aimag(s)

  do prec = 0, 1
    call monolis_hip_prm_initialize(Iarray, Rarray)
    Iarray(1) = monolis_hip_iter_COCG
    Iarray(2) = prec
    Iarray(20) = 0
    Iarray(21) = 0
    Iarray(22) = 0
    Rarray(1) = 1.0d-12
    x = (0.0d0, 0.0d0)
    call monolis_hip_solve_C(N, ND, index, item, A, b, x, Iarray, Rarray)
    write(*,"(a,i2,a,i4,a,1pe10.3,a,1pe10.3)") "COCG prec", prec, " iter", Iarray(4), &
      " resid", Rarray(2), " max|x-(1+i)|", maxval(abs(x - (1.0d0, 1.0d0)))
    if (Iarray(5) /= 0 .or. Iarray(4) < 1) ok = .false.
    if (maxval(abs(x - (1.0d0, 1.0d0))) > 1.0d-9) ok = .false.
  enddo

  if (ok) then
    write(*,"(a)") "FORTRAN_COCG_OK"
  else
    stop 1
  endif
contains
  !> deterministic values in [0, 1): a fixed linear congruential sequence
  real(c_double) function u01(state)
    integer(c_int), intent(inout) :: state
    state = mod(state * 1103 + 12345, 65536)
    u01 = real(state, c_double) / 65536.0d0
  end function u01
end program test_cocg_chain
