!> ISO_C_BINDING interface to libmonolis_hip.so for Fortran callers.
!>
!> The bind(c) interfaces are the flat entry points monolis itself exports
!> (wrapper/solver/solver_wrapper.f90:11-92, wrapper/linalg/matvec_wrap.f90:13-108,
!> wrapper/linalg/inner_product_wrap.f90:35-55), so a Fortran program that owns
!> monolis's BCSR arrays (index 0-based offsets, item 1-based columns, A row-major
!> blocks) calls the GPU path directly.  monolis_hip_solve_R / _matvec_R / _inner_product_R
!> fill in the one-rank communication arguments (comm_size 1, empty halo tables).
!> Iarray / Rarray are indexed as in monolis (Iarray(1) = method, Rarray(1) = tol);
!> in C they are the same arrays offset by one element (C slot k == Fortran slot k),
!> so they are passed starting at element 0, exactly as solver_wrapper.f90 copies them.
module mod_monolis_hip
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: monolis_hip_iter_CG = 1
  integer(c_int), parameter :: monolis_hip_iter_GropCG = 2
  integer(c_int), parameter :: monolis_hip_iter_PipeCG = 3
  integer(c_int), parameter :: monolis_hip_iter_PipeCR = 4
  integer(c_int), parameter :: monolis_hip_iter_BiCGSTAB = 5
  integer(c_int), parameter :: monolis_hip_iter_PipeBiCGSTAB = 6
  integer(c_int), parameter :: monolis_hip_iter_BiCGSTAB_noprec = 7
  integer(c_int), parameter :: monolis_hip_iter_PipeBiCGSTAB_noprec = 8
  integer(c_int), parameter :: monolis_hip_iter_COCG = 13
  integer(c_int), parameter :: monolis_hip_prec_NONE = 0
  integer(c_int), parameter :: monolis_hip_prec_DIAG = 1

  interface
    subroutine monolis_solve_R_c_main(n, np, nz, n_dof, NPNDOF, NZNDOF2, &
        n_dof_list, n_dof_index, n_dof_index2, A, x, b, index, item, &
        my_rank, comm, comm_size, &
        recv_n_neib, recv_nitem, recv_neib_pe, recv_index, recv_item, &
        send_n_neib, send_nitem, send_neib_pe, send_index, send_item, &
        Iarray, Rarray) bind(c, name="monolis_solve_R_c_main")
      import :: c_int, c_double
      integer(c_int), value :: n, np, nz, n_dof, NPNDOF, NZNDOF2
      integer(c_int) :: n_dof_list(*), n_dof_index(*), n_dof_index2(*)
      real(c_double) :: A(*), x(*), b(*)
      integer(c_int) :: index(*), item(*)
      integer(c_int), value :: my_rank, comm, comm_size
      integer(c_int), value :: recv_n_neib, recv_nitem
      integer(c_int) :: recv_neib_pe(*), recv_index(*), recv_item(*)
      integer(c_int), value :: send_n_neib, send_nitem
      integer(c_int) :: send_neib_pe(*), send_index(*), send_item(*)
      integer(c_int) :: Iarray(0:99)
      real(c_double) :: Rarray(0:99)
    end subroutine monolis_solve_R_c_main

    subroutine monolis_matvec_product_R_c_main(n, np, nz, n_dof, NPNDOF, NZNDOF2, &
        n_dof_list, n_dof_index, n_dof_index2, A, x, y, index, item, &
        my_rank, comm, comm_size, &
        recv_n_neib, recv_nitem, recv_neib_pe, recv_index, recv_item, &
        send_n_neib, send_nitem, send_neib_pe, send_index, send_item) &
        bind(c, name="monolis_matvec_product_R_c_main")
      import :: c_int, c_double
      integer(c_int), value :: n, np, nz, n_dof, NPNDOF, NZNDOF2
      integer(c_int) :: n_dof_list(*), n_dof_index(*), n_dof_index2(*)
      real(c_double) :: A(*), x(*), y(*)
      integer(c_int) :: index(*), item(*)
      integer(c_int), value :: my_rank, comm, comm_size
      integer(c_int), value :: recv_n_neib, recv_nitem
      integer(c_int) :: recv_neib_pe(*), recv_index(*), recv_item(*)
      integer(c_int), value :: send_n_neib, send_nitem
      integer(c_int) :: send_neib_pe(*), send_index(*), send_item(*)
    end subroutine monolis_matvec_product_R_c_main

    subroutine monolis_inner_product_R_c_main(m, x, y, sum, comm) &
        bind(c, name="monolis_inner_product_R_c_main")
      import :: c_int, c_double
      integer(c_int), value :: m
      real(c_double) :: x(*), y(*)
      real(c_double) :: sum
      integer(c_int), value :: comm
    end subroutine monolis_inner_product_R_c_main

    ! complex path (wrapper/solver/solver_wrapper.f90:94-175, wrapper/linalg/matvec_wrap.f90:112-207,
    ! wrapper/linalg/inner_product_wrap.f90:59-77)
    subroutine monolis_solve_C_c_main(n, np, nz, n_dof, NPNDOF, NZNDOF2, &
        n_dof_list, n_dof_index, n_dof_index2, A, x, b, index, item, &
        my_rank, comm, comm_size, &
        recv_n_neib, recv_nitem, recv_neib_pe, recv_index, recv_item, &
        send_n_neib, send_nitem, send_neib_pe, send_index, send_item, &
        Iarray, Rarray) bind(c, name="monolis_solve_C_c_main")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: n, np, nz, n_dof, NPNDOF, NZNDOF2
      integer(c_int) :: n_dof_list(*), n_dof_index(*), n_dof_index2(*)
      complex(c_double_complex) :: A(*), x(*), b(*)
      integer(c_int) :: index(*), item(*)
      integer(c_int), value :: my_rank, comm, comm_size
      integer(c_int), value :: recv_n_neib, recv_nitem
      integer(c_int) :: recv_neib_pe(*), recv_index(*), recv_item(*)
      integer(c_int), value :: send_n_neib, send_nitem
      integer(c_int) :: send_neib_pe(*), send_index(*), send_item(*)
      integer(c_int) :: Iarray(0:99)
      real(c_double) :: Rarray(0:99)
    end subroutine monolis_solve_C_c_main

    subroutine monolis_matvec_product_C_c_main(n, np, nz, n_dof, NPNDOF, NZNDOF2, &
        n_dof_list, n_dof_index, n_dof_index2, A, x, y, index, item, &
        my_rank, comm, comm_size, &
        recv_n_neib, recv_nitem, recv_neib_pe, recv_index, recv_item, &
        send_n_neib, send_nitem, send_neib_pe, send_index, send_item) &
        bind(c, name="monolis_matvec_product_C_c_main")
      import :: c_int, c_double_complex
      integer(c_int), value :: n, np, nz, n_dof, NPNDOF, NZNDOF2
      integer(c_int) :: n_dof_list(*), n_dof_index(*), n_dof_index2(*)
      complex(c_double_complex) :: A(*), x(*), y(*)
      integer(c_int) :: index(*), item(*)
      integer(c_int), value :: my_rank, comm, comm_size
      integer(c_int), value :: recv_n_neib, recv_nitem
      integer(c_int) :: recv_neib_pe(*), recv_index(*), recv_item(*)
      integer(c_int), value :: send_n_neib, send_nitem
      integer(c_int) :: send_neib_pe(*), send_index(*), send_item(*)
    end subroutine monolis_matvec_product_C_c_main

    subroutine monolis_inner_product_C_c_main(m, x, y, sum, comm) &
        bind(c, name="monolis_inner_product_C_c_main")
      import :: c_int, c_double_complex
      integer(c_int), value :: m
      complex(c_double_complex) :: x(*), y(*)
      complex(c_double_complex) :: sum
      integer(c_int), value :: comm
    end subroutine monolis_inner_product_C_c_main
  end interface

contains

  !> monolis_prm_initialize defaults (src/define/def_solver_prm.f90:183-222)
  subroutine monolis_hip_prm_initialize(Iarray, Rarray)
    integer(c_int), intent(out) :: Iarray(100)
    real(c_double), intent(out) :: Rarray(100)
    Iarray = 0
    Rarray = 0.0d0
    Iarray(1) = 1        ! method: CG
    Iarray(2) = 1        ! precond: DIAG
    Iarray(3) = 1000     ! max_iter
    Iarray(5) = -1       ! ierr
    Iarray(10) = 1       ! is_init_x
    Iarray(16) = 1       ! is_error_abort
    Iarray(20) = 1       ! show_iterlog
    Iarray(21) = 1       ! show_time
    Iarray(22) = 1       ! show_summary
    Iarray(24) = 4       ! IDRS dim
    Rarray(1) = 1.0d-8   ! tol
  end subroutine monolis_hip_prm_initialize

  !> One-rank solve: X (in/out, N*ndof), B (in), Iarray/Rarray monolis parameter slots.
  subroutine monolis_hip_solve_R(N, ndof, index, item, A, B, X, Iarray, Rarray)
    integer(c_int), intent(in) :: N, ndof
    integer(c_int), intent(in) :: index(N + 1), item(*)
    real(c_double), intent(in) :: A(*), B(*)
    real(c_double), intent(inout) :: X(*)
    integer(c_int), intent(inout) :: Iarray(100)
    real(c_double), intent(inout) :: Rarray(100)
    integer(c_int) :: empty(1), idx(N + 1), itm(max(1, index(N + 1)))
    integer(c_int) :: Ic(0:99)
    real(c_double) :: Rc(0:99)
    integer(c_int) :: nz
    nz = index(N + 1)
    empty = 0
    idx = index
    itm(1:nz) = item(1:nz)
    ! C slot k == Fortran slot k: shift by one element (solver_wrapper.f90:69-75)
    Ic(0) = 0
    Ic(1:99) = Iarray(1:99)
    Rc(0) = 0.0d0
    Rc(1:99) = Rarray(1:99)
    call monolis_solve_R_c_main(N, N, nz, ndof, N*ndof, nz*ndof*ndof, empty, empty, empty, &
      A, X, B, idx, itm, 0, 0, 1, 0, 0, empty, empty, empty, 0, 0, empty, empty, empty, Ic, Rc)
    Iarray(1:99) = Ic(1:99)
    Rarray(1:99) = Rc(1:99)
  end subroutine monolis_hip_solve_R

  !> One-rank SpMV: Y(1:N*ndof) = A X.
  subroutine monolis_hip_matvec_R(N, ndof, index, item, A, X, Y)
    integer(c_int), intent(in) :: N, ndof
    integer(c_int), intent(in) :: index(N + 1), item(*)
    real(c_double), intent(in) :: A(*), X(*)
    real(c_double), intent(out) :: Y(*)
    integer(c_int) :: empty(1), idx(N + 1), itm(max(1, index(N + 1)))
    integer(c_int) :: nz
    nz = index(N + 1)
    empty = 0
    idx = index
    itm(1:nz) = item(1:nz)
    call monolis_matvec_product_R_c_main(N, N, nz, ndof, N*ndof, nz*ndof*ndof, empty, empty, empty, &
      A, X, Y, idx, itm, 0, 0, 1, 0, 0, empty, empty, empty, 0, 0, empty, empty, empty)
  end subroutine monolis_hip_matvec_R

  !> sum = sum_i X(i) Y(i) over m entries.
  subroutine monolis_hip_inner_product_R(m, X, Y, sum)
    integer(c_int), intent(in) :: m
    real(c_double), intent(in) :: X(*), Y(*)
    real(c_double), intent(out) :: sum
    call monolis_inner_product_R_c_main(m, X, Y, sum, 0)
  end subroutine monolis_hip_inner_product_R
  !> One-rank complex solve (COCG, Iarray(1) = 13): X (in/out, N*ndof), B (in).
  subroutine monolis_hip_solve_C(N, ndof, index, item, A, B, X, Iarray, Rarray)
    integer(c_int), intent(in) :: N, ndof
    integer(c_int), intent(in) :: index(N + 1), item(*)
    complex(c_double_complex), intent(in) :: A(*), B(*)
    complex(c_double_complex), intent(inout) :: X(*)
    integer(c_int), intent(inout) :: Iarray(100)
    real(c_double), intent(inout) :: Rarray(100)
    integer(c_int) :: empty(1), idx(N + 1), itm(max(1, index(N + 1)))
    integer(c_int) :: Ic(0:99)
    real(c_double) :: Rc(0:99)
    integer(c_int) :: nz
    nz = index(N + 1)
    empty = 0
    idx = index
    itm(1:nz) = item(1:nz)
    Ic(0) = 0
    Ic(1:99) = Iarray(1:99)
    Rc(0) = 0.0d0
    Rc(1:99) = Rarray(1:99)
    call monolis_solve_C_c_main(N, N, nz, ndof, N*ndof, nz*ndof*ndof, empty, empty, empty, &
      A, X, B, idx, itm, 0, 0, 1, 0, 0, empty, empty, empty, 0, 0, empty, empty, empty, Ic, Rc)
    Iarray(1:99) = Ic(1:99)
    Rarray(1:99) = Rc(1:99)
  end subroutine monolis_hip_solve_C

  !> One-rank complex SpMV: Y(1:N*ndof) = A X.
  subroutine monolis_hip_matvec_C(N, ndof, index, item, A, X, Y)
    integer(c_int), intent(in) :: N, ndof
    integer(c_int), intent(in) :: index(N + 1), item(*)
    complex(c_double_complex), intent(in) :: A(*), X(*)
    complex(c_double_complex), intent(out) :: Y(*)
    integer(c_int) :: empty(1), idx(N + 1), itm(max(1, index(N + 1)))
    integer(c_int) :: nz
    nz = index(N + 1)
    empty = 0
    idx = index
    itm(1:nz) = item(1:nz)
    call monolis_matvec_product_C_c_main(N, N, nz, ndof, N*ndof, nz*ndof*ndof, empty, empty, empty, &
      A, X, Y, idx, itm, 0, 0, 1, 0, 0, empty, empty, empty, 0, 0, empty, empty, empty)
  end subroutine monolis_hip_matvec_C

  !> sum = sum_i X(i) Y(i) over m entries, unconjugated.
  subroutine monolis_hip_inner_product_C(m, X, Y, sum)
    integer(c_int), intent(in) :: m
    complex(c_double_complex), intent(in) :: X(*), Y(*)
    complex(c_double_complex), intent(out) :: sum
    call monolis_inner_product_C_c_main(m, X, Y, sum, 0)
  end subroutine monolis_hip_inner_product_C
end module mod_monolis_hip
