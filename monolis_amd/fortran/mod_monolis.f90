!> monolis's Fortran API over the GPU path (source-level drop-in).
!>
!> The subroutines keep monolis's names and argument lists
!> (src/define/def_struc.f90, def_solver_prm_util.f90, src/matrix/spmat_nzpattern.f90,
!> spmat_handler.f90, src/solver/solver.f90:29-49, src/linalg/matvec.f90:15-33,
!> inner_product.f90:104-125) with 1-based node / DOF indices, and call the C API
!> of include/monolis.h (libmonolis_front.so), which converts them and runs the
!> solves and products on the GPU.  monolis_structure and monolis_com are
!> bind(c) images of the C structs, so monolis%prm%Iarray(k), monolis%mat%N,
!> monolis%mat%NP and monolis%mat%NDOF read as in monolis; the matrix arrays are
!> C pointers (c_f_pointer to view them).
module mod_monolis
  use iso_c_binding
  implicit none

  integer, parameter :: kint = c_int
  integer, parameter :: kdouble = c_double

  integer(kint), parameter :: monolis_iter_CG = 1
  integer(kint), parameter :: monolis_iter_GropCG = 2
  integer(kint), parameter :: monolis_iter_PipeCG = 3
  integer(kint), parameter :: monolis_iter_PipeCR = 4
  integer(kint), parameter :: monolis_iter_BiCGSTAB = 5
  integer(kint), parameter :: monolis_iter_PipeBiCGSTAB = 6
  integer(kint), parameter :: monolis_iter_BiCGSTAB_noprec = 7
  integer(kint), parameter :: monolis_iter_PipeBiCGSTAB_noprec = 8
  integer(kint), parameter :: monolis_iter_COCG = 13
  integer(kint), parameter :: monolis_prec_NONE = 0
  integer(kint), parameter :: monolis_prec_DIAG = 1
  integer(kint), parameter :: monolis_prec_SOR = 2
  integer(kint), parameter :: monolis_prec_LU = 3

  !> MONOLIS_PRM (monolis_def_solver_prm_c.h:136-142); Iarray(k) / Rarray(k) are monolis's slots k
  type, bind(c) :: monolis_prm
    character(kind=c_char) :: com_top_dir_name(1024)
    character(kind=c_char) :: com_part_dir_name(1024)
    character(kind=c_char) :: com_file_name(1024)
    integer(c_int) :: Iarray(0:99)
    real(c_double) :: Rarray(0:99)
  end type monolis_prm

  type, bind(c) :: monolis_mat_val
    type(c_ptr) :: A, U, D, L, X, B
  end type monolis_mat_val

  type, bind(c) :: monolis_mat_scsr
    type(c_ptr) :: indexU, indexL, itemU, itemL
  end type monolis_mat_scsr

  type, bind(c) :: monolis_mat_csr
    type(c_ptr) :: index, item
  end type monolis_mat_csr

  type, bind(c) :: monolis_mat_csc
    type(c_ptr) :: index, item, perm
  end type monolis_mat_csc

  !> MONOLIS_MAT (monolis_def_mat_c.h:52-117)
  type, bind(c) :: monolis_mat
    integer(c_int) :: N, NP, NDOF
    type(c_ptr) :: n_dof_list, n_dof_index, n_dof_index2
    type(monolis_mat_val) :: R
    type(monolis_mat_val) :: C
    type(monolis_mat_scsr) :: SCSR
    type(monolis_mat_csr) :: CSR
    type(monolis_mat_csc) :: CSC
  end type monolis_mat

  !> MONOLIS (monolis_def_struc_c.h:17-21)
  type, bind(c) :: monolis_structure
    type(monolis_prm) :: prm
    type(monolis_mat) :: mat
    type(monolis_mat) :: prec
  end type monolis_structure

  !> MONOLIS_COM of include/monolis.h (one rank)
  type, bind(c) :: monolis_com
    integer(c_int) :: my_rank, comm, comm_size, n_internal_vertex
    integer(c_int) :: recv_n_neib
    type(c_ptr) :: recv_neib_pe, recv_index, recv_item
    integer(c_int) :: send_n_neib
    type(c_ptr) :: send_neib_pe, send_index, send_item
  end type monolis_com

  interface
    subroutine c_global_initialize() bind(c, name="monolis_global_initialize")
    end subroutine
    subroutine c_global_finalize() bind(c, name="monolis_global_finalize")
    end subroutine
    subroutine c_initialize(mat) bind(c, name="monolis_initialize")
      import :: monolis_structure
      type(monolis_structure) :: mat
    end subroutine
    subroutine c_finalize(mat) bind(c, name="monolis_finalize")
      import :: monolis_structure
      type(monolis_structure) :: mat
    end subroutine
    subroutine c_com_initialize_by_self(com) bind(c, name="monolis_com_initialize_by_self")
      import :: monolis_com
      type(monolis_com) :: com
    end subroutine
    subroutine c_com_finalize(com) bind(c, name="monolis_com_finalize")
      import :: monolis_com
      type(monolis_com) :: com
    end subroutine
    subroutine c_pattern_simple_mesh_R(mat, n_node, n_base, n_dof, n_elem, elem) &
        bind(c, name="monolis_get_nonzero_pattern_by_simple_mesh_R")
      import :: monolis_structure, c_int, c_ptr
      type(monolis_structure) :: mat
      integer(c_int), value :: n_node, n_base, n_dof, n_elem
      type(c_ptr) :: elem(*)
    end subroutine
    subroutine c_pattern_simple_mesh_C(mat, n_node, n_base, n_dof, n_elem, elem) &
        bind(c, name="monolis_get_nonzero_pattern_by_simple_mesh_C")
      import :: monolis_structure, c_int, c_ptr
      type(monolis_structure) :: mat
      integer(c_int), value :: n_node, n_base, n_dof, n_elem
      type(c_ptr) :: elem(*)
    end subroutine
    subroutine c_pattern_simple_mesh_V_R(mat, n_node, n_base, n_dof_list, n_elem, elem) &
        bind(c, name="monolis_get_nonzero_pattern_by_simple_mesh_V_R")
      import :: monolis_structure, c_int, c_ptr
      type(monolis_structure) :: mat
      integer(c_int), value :: n_node, n_base, n_elem
      integer(c_int) :: n_dof_list(*)
      type(c_ptr) :: elem(*)
    end subroutine
    subroutine c_pattern_graph_R(mat, n_node, n_dof, index, item) &
        bind(c, name="monolis_get_nonzero_pattern_by_nodal_graph_R")
      import :: monolis_structure, c_int
      type(monolis_structure) :: mat
      integer(c_int), value :: n_node, n_dof
      integer(c_int) :: index(*), item(*)
    end subroutine
    subroutine c_set_scalar_R(mat, i, j, si, sj, val) bind(c, name="monolis_set_scalar_to_sparse_matrix_R")
      import :: monolis_structure, c_int, c_double
      type(monolis_structure) :: mat
      integer(c_int), value :: i, j, si, sj
      real(c_double), value :: val
    end subroutine
    subroutine c_add_scalar_R(mat, i, j, si, sj, val) bind(c, name="monolis_add_scalar_to_sparse_matrix_R")
      import :: monolis_structure, c_int, c_double
      type(monolis_structure) :: mat
      integer(c_int), value :: i, j, si, sj
      real(c_double), value :: val
    end subroutine
    subroutine c_get_scalar_R(mat, i, j, si, sj, val, is_find) bind(c, name="monolis_get_scalar_from_sparse_matrix_R")
      import :: monolis_structure, c_int, c_double, c_bool
      type(monolis_structure) :: mat
      integer(c_int), value :: i, j, si, sj
      real(c_double) :: val
      logical(c_bool) :: is_find
    end subroutine
    subroutine c_add_scalar_C(mat, i, j, si, sj, val) bind(c, name="monolis_add_scalar_to_sparse_matrix_C")
      import :: monolis_structure, c_int, c_double_complex
      type(monolis_structure) :: mat
      integer(c_int), value :: i, j, si, sj
      complex(c_double_complex), value :: val
    end subroutine
    subroutine c_add_matrix_R(mat, n_base, conn, val) bind(c, name="monolis_add_matrix_to_sparse_matrix_R")
      import :: monolis_structure, c_int, c_ptr
      type(monolis_structure) :: mat
      integer(c_int), value :: n_base
      integer(c_int) :: conn(*)
      type(c_ptr) :: val(*)
    end subroutine
    subroutine c_dirichlet_R(mat, b, node_id, ndof_bc, val) bind(c, name="monolis_set_Dirichlet_bc_R")
      import :: monolis_structure, c_int, c_double
      type(monolis_structure) :: mat
      real(c_double) :: b(*)
      integer(c_int), value :: node_id, ndof_bc
      real(c_double), value :: val
    end subroutine
    subroutine c_solve_R(mat, com, b, x) bind(c, name="monolis_solve_R")
      import :: monolis_structure, monolis_com, c_double
      type(monolis_structure) :: mat
      type(monolis_com) :: com
      real(c_double) :: b(*), x(*)
    end subroutine
    subroutine c_solve_C(mat, com, b, x) bind(c, name="monolis_solve_C")
      import :: monolis_structure, monolis_com, c_double_complex
      type(monolis_structure) :: mat
      type(monolis_com) :: com
      complex(c_double_complex) :: b(*), x(*)
    end subroutine
    subroutine c_matvec_R(mat, com, x, y) bind(c, name="monolis_matvec_product_R")
      import :: monolis_structure, monolis_com, c_double
      type(monolis_structure) :: mat
      type(monolis_com) :: com
      real(c_double) :: x(*), y(*)
    end subroutine
    subroutine c_matvec_C(mat, com, x, y) bind(c, name="monolis_matvec_product_C")
      import :: monolis_structure, monolis_com, c_double_complex
      type(monolis_structure) :: mat
      type(monolis_com) :: com
      complex(c_double_complex) :: x(*), y(*)
    end subroutine
    subroutine c_inner_product_R(mat, com, n_dof, x, y, sum) bind(c, name="monolis_inner_product_R")
      import :: monolis_structure, monolis_com, c_int, c_double
      type(monolis_structure) :: mat
      type(monolis_com) :: com
      integer(c_int), value :: n_dof
      real(c_double) :: x(*), y(*), sum
    end subroutine
  end interface

contains

  ! ---- initialisation (src/define/def_struc.f90) ----------------------------
  subroutine monolis_global_initialize()
    call c_global_initialize()
  end subroutine monolis_global_initialize

  subroutine monolis_global_finalize()
    call c_global_finalize()
  end subroutine monolis_global_finalize

  subroutine monolis_initialize(monolis)
    type(monolis_structure), target :: monolis
    call c_initialize(monolis)
  end subroutine monolis_initialize

  subroutine monolis_finalize(monolis)
    type(monolis_structure), target :: monolis
    call c_finalize(monolis)
  end subroutine monolis_finalize

  subroutine monolis_com_initialize_by_self(com)
    type(monolis_com), target :: com
    call c_com_initialize_by_self(com)
  end subroutine monolis_com_initialize_by_self

  subroutine monolis_com_finalize(com)
    type(monolis_com), target :: com
    call c_com_finalize(com)
  end subroutine monolis_com_finalize

  ! ---- parameters (src/define/def_solver_prm_util.f90) ------------------------
  subroutine monolis_set_method(monolis, param)
    type(monolis_structure) :: monolis
    integer(kint), intent(in) :: param
    monolis%prm%Iarray(1) = param
  end subroutine monolis_set_method

  subroutine monolis_set_precond(monolis, param)
    type(monolis_structure) :: monolis
    integer(kint), intent(in) :: param
    monolis%prm%Iarray(2) = param
  end subroutine monolis_set_precond

  subroutine monolis_set_maxiter(monolis, param)
    type(monolis_structure) :: monolis
    integer(kint), intent(in) :: param
    monolis%prm%Iarray(3) = param
  end subroutine monolis_set_maxiter

  subroutine monolis_set_tolerance(monolis, val)
    type(monolis_structure) :: monolis
    real(kdouble), intent(in) :: val
    monolis%prm%Rarray(1) = val
  end subroutine monolis_set_tolerance

  subroutine set_flag(monolis, slot, param)
    type(monolis_structure) :: monolis
    integer, intent(in) :: slot
    logical, intent(in) :: param
    monolis%prm%Iarray(slot) = merge(1, 0, param)
  end subroutine set_flag

  subroutine monolis_set_init_x(monolis, param)
    type(monolis_structure) :: monolis
    logical, intent(in) :: param
    call set_flag(monolis, 10, param)
  end subroutine monolis_set_init_x

  subroutine monolis_set_error_abort(monolis, param)
    type(monolis_structure) :: monolis
    logical, intent(in) :: param
    call set_flag(monolis, 16, param)
  end subroutine monolis_set_error_abort

  subroutine monolis_set_performance_measurement(monolis, param)
    type(monolis_structure) :: monolis
    logical, intent(in) :: param
    call set_flag(monolis, 13, param)
  end subroutine monolis_set_performance_measurement

  subroutine monolis_show_iterlog(monolis, param)
    type(monolis_structure) :: monolis
    logical, intent(in) :: param
    call set_flag(monolis, 20, param)
  end subroutine monolis_show_iterlog

  subroutine monolis_show_timelog(monolis, param)
    type(monolis_structure) :: monolis
    logical, intent(in) :: param
    call set_flag(monolis, 21, param)
  end subroutine monolis_show_timelog

  subroutine monolis_show_summary(monolis, param)
    type(monolis_structure) :: monolis
    logical, intent(in) :: param
    call set_flag(monolis, 22, param)
  end subroutine monolis_show_summary

  subroutine monolis_show_timelog_statistics(monolis, param)
    type(monolis_structure) :: monolis
    logical, intent(in) :: param
    call set_flag(monolis, 23, param)
  end subroutine monolis_show_timelog_statistics

  subroutine monolis_get_converge_iter(monolis, param)
    type(monolis_structure) :: monolis
    integer(kint), intent(out) :: param
    param = monolis%prm%Iarray(4)
  end subroutine monolis_get_converge_iter

  subroutine monolis_get_error_tag(monolis, param)
    type(monolis_structure) :: monolis
    integer(kint), intent(out) :: param
    param = monolis%prm%Iarray(5)
  end subroutine monolis_get_error_tag

  subroutine monolis_get_converge_residual(monolis, val)
    type(monolis_structure) :: monolis
    real(kdouble), intent(out) :: val
    val = monolis%prm%Rarray(2)
  end subroutine monolis_get_converge_residual

  subroutine monolis_get_time_solver(monolis, val)
    type(monolis_structure) :: monolis
    real(kdouble), intent(out) :: val
    val = monolis%prm%Rarray(3)
  end subroutine monolis_get_time_solver

  ! ---- nonzero patterns (src/matrix/spmat_nzpattern.f90) ----------------------
  ! elem(n_base, n_elem): 1-based node ids, as monolis's Fortran API takes them
  subroutine elem_rows(n_base, n_elem, elem, e0, rows)
    integer(kint), intent(in) :: n_base, n_elem
    integer(kint), intent(in) :: elem(:,:)
    integer(c_int), allocatable, target, intent(out) :: e0(:,:)
    type(c_ptr), allocatable, intent(out) :: rows(:)
    integer :: e
    allocate(e0(n_base, max(n_elem, 1)))
    allocate(rows(max(n_elem, 1)))
    do e = 1, n_elem
      e0(:, e) = elem(1:n_base, e) - 1
      rows(e) = c_loc(e0(1, e))
    enddo
  end subroutine elem_rows

  subroutine monolis_get_nonzero_pattern_by_simple_mesh_R(monolis, n_node, n_base, ndof, n_elem, elem)
    type(monolis_structure), target :: monolis
    integer(kint), intent(in) :: n_node, n_base, ndof, n_elem
    integer(kint), intent(in) :: elem(:,:)
    integer(c_int), allocatable, target :: e0(:,:)
    type(c_ptr), allocatable :: rows(:)
    call elem_rows(n_base, n_elem, elem, e0, rows)
    call c_pattern_simple_mesh_R(monolis, n_node, n_base, ndof, n_elem, rows)
  end subroutine monolis_get_nonzero_pattern_by_simple_mesh_R

  subroutine monolis_get_nonzero_pattern_by_simple_mesh_C(monolis, n_node, n_base, ndof, n_elem, elem)
    type(monolis_structure), target :: monolis
    integer(kint), intent(in) :: n_node, n_base, ndof, n_elem
    integer(kint), intent(in) :: elem(:,:)
    integer(c_int), allocatable, target :: e0(:,:)
    type(c_ptr), allocatable :: rows(:)
    call elem_rows(n_base, n_elem, elem, e0, rows)
    call c_pattern_simple_mesh_C(monolis, n_node, n_base, ndof, n_elem, rows)
  end subroutine monolis_get_nonzero_pattern_by_simple_mesh_C

  subroutine monolis_get_nonzero_pattern_by_simple_mesh_V_R(monolis, n_node, n_base, n_dof_list, n_elem, elem)
    type(monolis_structure), target :: monolis
    integer(kint), intent(in) :: n_node, n_base, n_elem
    integer(kint), intent(in) :: n_dof_list(:)
    integer(kint), intent(in) :: elem(:,:)
    integer(c_int), allocatable, target :: e0(:,:)
    type(c_ptr), allocatable :: rows(:)
    integer(c_int), allocatable :: lst(:)
    lst = n_dof_list(1:n_node)
    call elem_rows(n_base, n_elem, elem, e0, rows)
    call c_pattern_simple_mesh_V_R(monolis, n_node, n_base, lst, n_elem, rows)
  end subroutine monolis_get_nonzero_pattern_by_simple_mesh_V_R

  ! index(n_node+1): 0-based offsets; item: 1-based neighbour ids (gedatsu's graph)
  subroutine monolis_get_nonzero_pattern_by_nodal_graph_R(monolis, n_node, ndof, index, item)
    type(monolis_structure), target :: monolis
    integer(kint), intent(in) :: n_node, ndof
    integer(kint), intent(in) :: index(:), item(:)
    integer(c_int), allocatable :: it0(:), ix(:)
    ix = index(1:n_node + 1)
    allocate(it0(max(index(n_node + 1), 1)))
    if(index(n_node + 1) > 0) it0(1:index(n_node + 1)) = item(1:index(n_node + 1)) - 1
    call c_pattern_graph_R(monolis, n_node, ndof, ix, it0)
  end subroutine monolis_get_nonzero_pattern_by_nodal_graph_R

  ! ---- assembly (src/matrix/spmat_handler.f90), 1-based indices ---------------
  subroutine monolis_set_scalar_to_sparse_matrix_R(monolis, i, j, sub_i, sub_j, val)
    type(monolis_structure), target :: monolis
    integer(kint), intent(in) :: i, j, sub_i, sub_j
    real(kdouble), intent(in) :: val
    call c_set_scalar_R(monolis, i - 1, j - 1, sub_i - 1, sub_j - 1, val)
  end subroutine monolis_set_scalar_to_sparse_matrix_R

  subroutine monolis_add_scalar_to_sparse_matrix_R(monolis, i, j, sub_i, sub_j, val)
    type(monolis_structure), target :: monolis
    integer(kint), intent(in) :: i, j, sub_i, sub_j
    real(kdouble), intent(in) :: val
    call c_add_scalar_R(monolis, i - 1, j - 1, sub_i - 1, sub_j - 1, val)
  end subroutine monolis_add_scalar_to_sparse_matrix_R

  subroutine monolis_get_scalar_from_sparse_matrix_R(monolis, i, j, sub_i, sub_j, val, is_find)
    type(monolis_structure), target :: monolis
    integer(kint), intent(in) :: i, j, sub_i, sub_j
    real(kdouble), intent(out) :: val
    logical, intent(out) :: is_find
    logical(c_bool) :: f
    call c_get_scalar_R(monolis, i - 1, j - 1, sub_i - 1, sub_j - 1, val, f)
    is_find = f
  end subroutine monolis_get_scalar_from_sparse_matrix_R

  subroutine monolis_add_scalar_to_sparse_matrix_C(monolis, i, j, sub_i, sub_j, val)
    type(monolis_structure), target :: monolis
    integer(kint), intent(in) :: i, j, sub_i, sub_j
    complex(kdouble), intent(in) :: val
    call c_add_scalar_C(monolis, i - 1, j - 1, sub_i - 1, sub_j - 1, val)
  end subroutine monolis_add_scalar_to_sparse_matrix_C

  ! val(n_base*ndof, n_base*ndof): row (a, k1), column (c, k2) of the element matrix
  subroutine monolis_add_matrix_to_sparse_matrix_R(monolis, n_base, connectivity, val)
    type(monolis_structure), target :: monolis
    integer(kint), intent(in) :: n_base
    integer(kint), intent(in) :: connectivity(:)
    real(kdouble), intent(in) :: val(:,:)
    real(c_double), allocatable, target :: rowmajor(:,:)
    type(c_ptr), allocatable :: rows(:)
    integer(c_int), allocatable :: conn0(:)
    integer :: r, nr
    nr = size(val, 1)
    allocate(rowmajor(size(val, 2), nr), rows(nr))
    rowmajor = transpose(val)
    do r = 1, nr
      rows(r) = c_loc(rowmajor(1, r))
    enddo
    conn0 = connectivity(1:n_base) - 1
    call c_add_matrix_R(monolis, n_base, conn0, rows)
  end subroutine monolis_add_matrix_to_sparse_matrix_R

  subroutine monolis_set_Dirichlet_bc_R(monolis, B, node_id, ndof_bc, val)
    type(monolis_structure), target :: monolis
    real(kdouble) :: B(:)
    integer(kint), intent(in) :: node_id, ndof_bc
    real(kdouble), intent(in) :: val
    call c_dirichlet_R(monolis, B, node_id - 1, ndof_bc - 1, val)
  end subroutine monolis_set_Dirichlet_bc_R

  ! ---- solver and linear algebra on the GPU path ------------------------------
  subroutine monolis_solve_R(monolis, com, B, X)
    type(monolis_structure), target :: monolis
    type(monolis_com), target :: com
    real(kdouble) :: B(:), X(:)
    call c_solve_R(monolis, com, B, X)
  end subroutine monolis_solve_R

  subroutine monolis_solve_C(monolis, com, B, X)
    type(monolis_structure), target :: monolis
    type(monolis_com), target :: com
    complex(kdouble) :: B(:), X(:)
    call c_solve_C(monolis, com, B, X)
  end subroutine monolis_solve_C

  subroutine monolis_matvec_product_R(monolis, com, X, Y)
    type(monolis_structure), target :: monolis
    type(monolis_com), target :: com
    real(kdouble) :: X(:), Y(:)
    call c_matvec_R(monolis, com, X, Y)
  end subroutine monolis_matvec_product_R

  subroutine monolis_matvec_product_C(monolis, com, X, Y)
    type(monolis_structure), target :: monolis
    type(monolis_com), target :: com
    complex(kdouble) :: X(:), Y(:)
    call c_matvec_C(monolis, com, X, Y)
  end subroutine monolis_matvec_product_C

  subroutine monolis_inner_product_R(monolis, com, ndof, X, Y, sum)
    type(monolis_structure), target :: monolis
    type(monolis_com), target :: com
    integer(kint), intent(in) :: ndof
    real(kdouble) :: X(:), Y(:)
    real(kdouble), intent(out) :: sum
    call c_inner_product_R(monolis, com, ndof, X, Y, sum)
  end subroutine monolis_inner_product_R

end module mod_monolis
