!> A Fortran program written against monolis's Fortran API (mod_monolis), the
!> shape of src_test/iterative/CG_test.f90:21-90 and BiCGSTAB_test.f90: a
!> 10-node chain from monolis_get_nonzero_pattern_by_simple_mesh_R, random
!> symmetric values through monolis_add_scalar_to_sparse_matrix_R, b = A 1 by
!> monolis_matvec_product_R, then monolis_solve_R must return x = 1 for CG,
!> BiCGSTAB and PipeCR with NONE and DIAG; the complex chain of COCG_test.f90
!> returns x = 1 + i; the Dirichlet values of spmat_handler_test.f90:119-132
!> on the host.  Prints FORTRAN_API_OK when every check passes.
program test_fortran_api
  use mod_monolis
  implicit none
  integer :: nfail, n_dof, prec, im, i
  integer, parameter :: methods(3) = [monolis_iter_CG, monolis_iter_BiCGSTAB, monolis_iter_PipeCR]

  nfail = 0
  call monolis_global_initialize()
  call dirichlet_check()
  do n_dof = 1, 3
    do prec = monolis_prec_NONE, monolis_prec_DIAG
      do im = 1, size(methods)
        call chain_solve(n_dof, prec, methods(im))
      enddo
    enddo
  enddo
  do i = 0, 1
    call cocg_chain(i)
  enddo
  call monolis_global_finalize()
  if(nfail == 0)then
    write(*,"(a)") "FORTRAN_API_OK"
  else
    write(*,"(a,i0)") "FORTRAN_API failures: ", nfail
    stop 1
  endif

contains

  subroutine chain_elems(elem)
    integer(kint), intent(out) :: elem(2, 9)
    integer :: k
    do k = 1, 9
      elem(1, k) = k
      elem(2, k) = k + 1
    enddo
  end subroutine chain_elems

  subroutine dirichlet_check()
    type(monolis_structure) :: mat
    integer(kint) :: elem(2, 3)
    real(kdouble) :: B(8), v
    real(kdouble), pointer :: A(:)
    logical :: found
    elem(:, 1) = [1, 2]
    elem(:, 2) = [2, 3]
    elem(:, 3) = [3, 4]
    call monolis_initialize(mat)
    call monolis_get_nonzero_pattern_by_simple_mesh_R(mat, 4, 2, 2, 3, elem)
    if(mat%mat%N /= 4 .or. mat%mat%NDOF /= 2) call fail("pattern sizes")
    call c_f_pointer(mat%mat%R%A, A, [40])
    A = 2.0d0
    B = 0.0d0
    call monolis_set_Dirichlet_bc_R(mat, B, 2, 1, 1.0d0)
    if(any(B /= [-2.0d0, -2.0d0, 1.0d0, -2.0d0, -2.0d0, -2.0d0, 0.0d0, 0.0d0])) call fail("Dirichlet B")
    call monolis_get_scalar_from_sparse_matrix_R(mat, 2, 2, 1, 1, v, found)
    if(.not. found .or. v /= 1.0d0) call fail("Dirichlet diagonal")
    call monolis_get_scalar_from_sparse_matrix_R(mat, 1, 4, 1, 1, v, found)
    if(found) call fail("block outside the pattern found")
    call monolis_finalize(mat)
  end subroutine dirichlet_check

  subroutine chain_solve(n_dof, prec, method)
    integer, intent(in) :: n_dof, prec, method
    type(monolis_structure) :: mat
    type(monolis_com) :: com
    integer(kint) :: elem(2, 9), i1, i2, j2, it
    real(kdouble) :: val, a(n_dof*10), b(n_dof*10), res
    call monolis_initialize(mat)
    call monolis_com_initialize_by_self(com)
    call chain_elems(elem)
    call monolis_get_nonzero_pattern_by_simple_mesh_R(mat, 10, 2, n_dof, 9, elem)
    do i1 = 1, 10
      do i2 = 1, n_dof
        call random_number(val)
        call monolis_add_scalar_to_sparse_matrix_R(mat, i1, i1, i2, i2, val + 2.0d0)
      enddo
    enddo
    do i1 = 1, 9
      do i2 = 1, n_dof
        do j2 = 1, n_dof
          call random_number(val)
          call monolis_add_scalar_to_sparse_matrix_R(mat, elem(1, i1), elem(2, i1), i2, j2, val)
          call monolis_add_scalar_to_sparse_matrix_R(mat, elem(2, i1), elem(1, i1), j2, i2, val)
        enddo
      enddo
    enddo
    a = 1.0d0
    call monolis_matvec_product_R(mat, com, a, b)
    call monolis_set_method(mat, method)
    call monolis_set_precond(mat, prec)
    call monolis_set_tolerance(mat, 1.0d-10)
    call monolis_show_iterlog(mat, .false.)
    call monolis_show_timelog(mat, .false.)
    call monolis_show_summary(mat, .false.)
    a = 0.0d0
    call monolis_solve_R(mat, com, b, a)
    call monolis_get_converge_iter(mat, it)
    call monolis_get_converge_residual(mat, res)
    if(maxval(abs(a - 1.0d0)) > 1.0d-6 .or. it < 1 .or. res > 1.0d-10)then
      write(*,"(a,3i3,i5,2es12.3)") "solve n_dof/prec/method, iter, resid, err:", n_dof, prec, method, it, res, &
        maxval(abs(a - 1.0d0))
      call fail("chain solve")
    endif
    call monolis_com_finalize(com)
    call monolis_finalize(mat)
  end subroutine chain_solve

  subroutine cocg_chain(prec)
    integer, intent(in) :: prec
    type(monolis_structure) :: mat
    type(monolis_com) :: com
    integer(kint) :: elem(2, 9), i1, i2, j2
    real(kdouble) :: vr, vi
    complex(kdouble) :: a(20), b(20)
    call monolis_initialize(mat)
    call monolis_com_initialize_by_self(com)
    call chain_elems(elem)
    call monolis_get_nonzero_pattern_by_simple_mesh_C(mat, 10, 2, 2, 9, elem)
    do i1 = 1, 10
      do i2 = 1, 2
        call random_number(vr)
        call random_number(vi)
        call monolis_add_scalar_to_sparse_matrix_C(mat, i1, i1, i2, i2, cmplx(vr + 2.0d0, vi + 2.0d0, kdouble))
      enddo
    enddo
    do i1 = 1, 9
      do i2 = 1, 2
        do j2 = 1, 2
          call random_number(vr)
          call random_number(vi)
          call monolis_add_scalar_to_sparse_matrix_C(mat, elem(1, i1), elem(2, i1), i2, j2, cmplx(vr, vi, kdouble))
          call monolis_add_scalar_to_sparse_matrix_C(mat, elem(2, i1), elem(1, i1), j2, i2, cmplx(vr, vi, kdouble))
        enddo
      enddo
    enddo
    a = (1.0d0, 1.0d0)
    call monolis_matvec_product_C(mat, com, a, b)
    call monolis_set_method(mat, monolis_iter_COCG)
    call monolis_set_precond(mat, prec)
    call monolis_set_tolerance(mat, 1.0d-10)
    call monolis_show_iterlog(mat, .false.)
    call monolis_show_timelog(mat, .false.)
    call monolis_show_summary(mat, .false.)
    a = (0.0d0, 0.0d0)
    call monolis_solve_C(mat, com, b, a)
    if(maxval(abs(a - (1.0d0, 1.0d0))) > 1.0d-6) call fail("COCG chain")
    call monolis_com_finalize(com)
    call monolis_finalize(mat)
  end subroutine cocg_chain

  subroutine fail(what)
    character(*), intent(in) :: what
    write(*,"(2a)") "FAIL: ", what
    nfail = nfail + 1
  end subroutine fail

end program test_fortran_api
