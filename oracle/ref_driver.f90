!-------------------------------------------------------------------------------
! Test-only driver around the reference's own hecmw1 solver (FrontISTR 5.9).
!
! This file is part of the oracle: it is linked against objects compiled
! unmodified from the reference sources (see build_ref.py) into
! oracle/_ref/libhecmw_ref.so.  Only the test-suite and the CPU-baseline leg of
! bench.py load that library; the product never does.
!
! It exposes plain C entry points (bind(c)) that wrap a caller's hecmwST_matrix
! arrays (no copies: Fortran pointers are aimed at the caller's memory) and call
! the reference routines:
!   ref_solve_krylov  -> hecmw_solve_CG / hecmw_solve_BiCGSTAB
!                        (hecmw1/src/solver/iterative/hecmw_solver_CG.f90:19,
!                         hecmw_solver_BiCGSTAB.f90:16)
!   ref_solve_iterative -> hecmw_solve_iterative (hecmw_solver_Iterative.f90:13)
!   ref_ordering      -> hecmw_matrix_ordering_RCM + hecmw_matrix_ordering_MC
!                        (hecmw_matrix_ordering_CM.f90:65, _MC.f90:15)
!   ref_matvec        -> hecmw_matvec (hecmw_solver_las.f90:42)
!   ref_precond       -> hecmw_precond_setup + hecmw_precond_apply
!                        (hecmw_precond.f90:30,86)
!   ref_rel_resid     -> hecmw_rel_resid_L2 (hecmw_solver_las.f90:124)
!   ref_matvec_mpc    -> hecmw_matvec with the MPC flag (hecmw_TtmatTvec_33)
!   ref_set_mpc       -> binds hecMESH%mpc for the calls that follow
!   ref_mat_dump      -> hecmw_mat_dump + hecmw_mat_dump_solution
!                        (hecmw_matrix_dump.f90:31,345)
!-------------------------------------------------------------------------------
module hecmw_ref_driver
  use iso_c_binding
  use hecmw_util
  implicit none
  private

  type(hecmwST_local_mesh), save, target :: mesh
  type(hecmwST_matrix), save, target :: mat
  ! multipoint constraints bound by ref_set_mpc (0: none), kept for every later call
  integer(kind=kint), save :: nmpc = 0

contains

  subroutine bind_matrix(n, np, npl, npu, d, al, au, indexl, iteml, indexu, itemu, &
      b, x, iarray, rarray)
    integer(c_int), value :: n, np, npl, npu
    type(c_ptr), value :: d, al, au, indexl, iteml, indexu, itemu, b, x
    integer(c_int), intent(in) :: iarray(100)
    real(c_double), intent(in) :: rarray(100)
    integer(kind=kint), pointer :: tmp(:)

    mat%N = n
    mat%NP = np
    mat%NPL = npl
    mat%NPU = npu
    mat%NPA = 0
    mat%NDOF = 3
    call c_f_pointer(d, mat%D, [9*np])
    call c_f_pointer(al, mat%AL, [max(9*npl,1)])
    call c_f_pointer(au, mat%AU, [max(9*npu,1)])
    call c_f_pointer(indexl, tmp, [np+1])
    mat%indexL(0:np) => tmp
    call c_f_pointer(indexu, tmp, [np+1])
    mat%indexU(0:np) => tmp
    call c_f_pointer(iteml, mat%itemL, [max(npl,1)])
    call c_f_pointer(itemu, mat%itemU, [max(npu,1)])
    call c_f_pointer(b, mat%B, [3*np])
    call c_f_pointer(x, mat%X, [3*np])
    nullify(mat%A, mat%indexA, mat%itemA)
    mat%Iarray(:) = iarray(:)
    mat%Rarray(:) = rarray(:)

    mesh%nn_internal = n
    mesh%n_node = np
    mesh%n_dof = 3
    mesh%n_neighbor_pe = 0
    mesh%my_rank = 0
    mesh%PETOT = 1
    mesh%zero = 0
    mesh%mpc%n_mpc = nmpc
  end subroutine bind_matrix

  !> hecMESH%mpc for the following calls (hecmw_util_f.F90:97-104): n_mpc
  !> constraints, mpc_index(0:n_mpc), mpc_item/mpc_dof/mpc_val per term (the
  !> caller's memory; n_mpc = 0 clears).  With MPC_METHOD = 2 (Iarray 13) the
  !> reference's solvers then multiply by T^T A T (hecmw_TtmatTvec_33).
  subroutine ref_set_mpc(n_mpc, index, item, dof, val) bind(c, name='ref_set_mpc')
    integer(c_int), value :: n_mpc
    type(c_ptr), value :: index, item, dof, val
    integer(kind=kint), pointer :: tmp(:)
    integer(kind=kint) :: nt
    nmpc = n_mpc
    if (n_mpc <= 0) then
      nmpc = 0
      return
    endif
    call c_f_pointer(index, tmp, [n_mpc + 1])
    mesh%mpc%mpc_index(0:n_mpc) => tmp
    nt = tmp(n_mpc + 1)
    call c_f_pointer(item, mesh%mpc%mpc_item, [max(nt, 1)])
    call c_f_pointer(dof, mesh%mpc%mpc_dof, [max(nt, 1)])
    call c_f_pointer(val, mesh%mpc%mpc_val, [max(nt, 1)])
  end subroutine ref_set_mpc

  subroutine set_threads(nthreads)
    !$ use omp_lib
    integer(c_int), value :: nthreads
    !$ if (nthreads > 0) call omp_set_num_threads(nthreads)
  end subroutine set_threads

  !> One Krylov solve through the reference's CG or BiCGSTAB routine.
  !> out(1)=iterations, out(2)=final RESID, out(3)=Tset, out(4)=Tsol,
  !> out(5)=Tcomm, out(6)=matvec time, out(7)=precond time.
  !> Returns the reference's error code (0 on success).
  function ref_solve_krylov(n, np, npl, npu, d, al, au, indexl, iteml, indexu, itemu, &
      b, x, iarray, rarray, nthreads, iarray_out, out) result(ierr) bind(c, name='ref_solve_krylov')
    use hecmw_solver_CG
    use hecmw_solver_BiCGSTAB
    use hecmw_solver_las
    use hecmw_precond
    use hecmw_matrix_misc
    integer(c_int), value :: n, np, npl, npu, nthreads
    type(c_ptr), value :: d, al, au, indexl, iteml, indexu, itemu, b, x
    integer(c_int), intent(in) :: iarray(100)
    real(c_double), intent(in) :: rarray(100)
    integer(c_int), intent(out) :: iarray_out(100)
    real(c_double), intent(out) :: out(7)
    integer(c_int) :: ierr
    integer(kind=kint) :: iter, error
    real(kind=kreal) :: resid, tset, tsol, tcomm

    call set_threads(nthreads)
    call bind_matrix(n, np, npl, npu, d, al, au, indexl, iteml, indexu, itemu, b, x, iarray, rarray)
    call hecmw_matvec_clear_timer()
    call hecmw_precond_clear_timer()
    iter = 0; error = 0; resid = 0.d0; tset = 0.d0; tsol = 0.d0; tcomm = 0.d0
    select case (hecmw_mat_get_method(mat))
      case (1)
        call hecmw_solve_CG(mesh, mat, iter, resid, error, tset, tsol, tcomm)
      case (2)
        call hecmw_solve_BiCGSTAB(mesh, mat, iter, resid, error, tset, tsol, tcomm)
      case default
        error = 1001
    end select
    out(1) = dble(iter)
    out(2) = resid
    out(3) = tset
    out(4) = tsol
    out(5) = tcomm
    out(6) = hecmw_matvec_get_timer()
    out(7) = hecmw_precond_get_timer()
    iarray_out(:) = mat%Iarray(:)
    ierr = error
  end function ref_solve_krylov

  !> The full iterative dispatcher (parameters, recycle flags, retry loop,
  !> post-solve relative residual and convergence flags).  The zero-diagonal
  !> check of the dispatcher stops the process in the reference, so callers
  !> only pass systems with non-zero diagonals.
  function ref_solve_iterative(n, np, npl, npu, d, al, au, indexl, iteml, indexu, itemu, &
      b, x, iarray, rarray, nthreads, iarray_out, rarray_out) result(ierr) &
      bind(c, name='ref_solve_iterative')
    use hecmw_solver_iterative
    integer(c_int), value :: n, np, npl, npu, nthreads
    type(c_ptr), value :: d, al, au, indexl, iteml, indexu, itemu, b, x
    integer(c_int), intent(in) :: iarray(100)
    real(c_double), intent(in) :: rarray(100)
    integer(c_int), intent(out) :: iarray_out(100)
    real(c_double), intent(out) :: rarray_out(100)
    integer(c_int) :: ierr
    call set_threads(nthreads)
    call bind_matrix(n, np, npl, npu, d, al, au, indexl, iteml, indexu, itemu, b, x, iarray, rarray)
    call hecmw_solve_iterative(mesh, mat)
    iarray_out(:) = mat%Iarray(:)
    rarray_out(:) = mat%Rarray(:)
    ierr = 0
  end function ref_solve_iterative

  !> RCM followed by the capped greedy multicolouring, exactly as
  !> hecmw_precond_SSOR_33_setup calls them.  perm_rcm/iperm_rcm receive the
  !> RCM permutation, perm/iperm/colorindex(0:ncolor) the final one.
  subroutine ref_ordering(n, indexl, iteml, indexu, itemu, ncolor_in, perm_rcm, iperm_rcm, &
      ncolor, colorindex, perm, iperm) bind(c, name='ref_ordering')
    use m_hecmw_matrix_ordering_CM
    use m_hecmw_matrix_ordering_MC
    integer(c_int), value :: n, ncolor_in
    type(c_ptr), value :: indexl, iteml, indexu, itemu
    integer(c_int), intent(out) :: perm_rcm(n), iperm_rcm(n), perm(n), iperm(n)
    integer(c_int), intent(out) :: ncolor
    integer(c_int), intent(out) :: colorindex(0:n)
    integer(kind=kint), pointer :: il(:), iu(:), tl(:), tu(:), tmp(:)
    call c_f_pointer(indexl, tmp, [n+1])
    il(0:n) => tmp
    call c_f_pointer(indexu, tmp, [n+1])
    iu(0:n) => tmp
    call c_f_pointer(iteml, tl, [max(il(n),1)])
    call c_f_pointer(itemu, tu, [max(iu(n),1)])
    call hecmw_matrix_ordering_RCM(n, il, tl, iu, tu, perm_rcm, iperm_rcm)
    call hecmw_matrix_ordering_MC(n, il, tl, iu, tu, perm_rcm, ncolor_in, ncolor, &
      colorindex, perm, iperm)
  end subroutine ref_ordering

  !> y = A x through hecmw_matvec (internal rows only are written).
  subroutine ref_matvec(n, np, npl, npu, d, al, au, indexl, iteml, indexu, itemu, &
      xin, y, nthreads) bind(c, name='ref_matvec')
    use hecmw_solver_las
    integer(c_int), value :: n, np, npl, npu, nthreads
    type(c_ptr), value :: d, al, au, indexl, iteml, indexu, itemu, xin, y
    real(kind=kreal), pointer :: xv(:), yv(:)
    integer(c_int) :: ia(100)
    real(c_double) :: ra(100)
    ia = 0; ra = 0.d0
    call set_threads(nthreads)
    call bind_matrix(n, np, npl, npu, d, al, au, indexl, iteml, indexu, itemu, xin, y, ia, ra)
    call c_f_pointer(xin, xv, [3*np])
    call c_f_pointer(y, yv, [3*np])
    call hecmw_matvec(mesh, mat, xv, yv)
  end subroutine ref_matvec

  !> y = T^T A T x through hecmw_matvec with the MPC flag set (Iarray 83,
  !> hecmw_TtmatTvec_33, hecmw_solver_las_33.F90:572-586) and the constraints of ref_set_mpc.
  subroutine ref_matvec_mpc(n, np, npl, npu, d, al, au, indexl, iteml, indexu, itemu, &
      xin, y, nthreads) bind(c, name='ref_matvec_mpc')
    use hecmw_solver_las
    use hecmw_matrix_misc
    integer(c_int), value :: n, np, npl, npu, nthreads
    type(c_ptr), value :: d, al, au, indexl, iteml, indexu, itemu, xin, y
    real(kind=kreal), pointer :: xv(:), yv(:)
    integer(c_int) :: ia(100)
    real(c_double) :: ra(100)
    ia = 0; ra = 0.d0
    call set_threads(nthreads)
    call bind_matrix(n, np, npl, npu, d, al, au, indexl, iteml, indexu, itemu, xin, y, ia, ra)
    call hecmw_mat_set_flag_mpcmatvec(mat, 1)
    call c_f_pointer(xin, xv, [3*np])
    call c_f_pointer(y, yv, [3*np])
    call hecmw_matvec(mesh, mat, xv, yv)
  end subroutine ref_matvec_mpc

  !> z = M^-1 r with the preconditioner selected by iarray (PRECOND, ITERPREMAX,
  !> NCOLOR_IN, SIGMA_DIAG); a fresh setup is forced.
  subroutine ref_precond(n, np, npl, npu, d, al, au, indexl, iteml, indexu, itemu, &
      r, z, iarray, rarray, nthreads) bind(c, name='ref_precond')
    use hecmw_precond
    integer(c_int), value :: n, np, npl, npu, nthreads
    type(c_ptr), value :: d, al, au, indexl, iteml, indexu, itemu, r, z
    integer(c_int), intent(in) :: iarray(100)
    real(c_double), intent(in) :: rarray(100)
    real(kind=kreal), pointer :: rv(:), zv(:)
    real(kind=kreal), allocatable :: wk(:)
    real(kind=kreal) :: tcomm
    call set_threads(nthreads)
    call bind_matrix(n, np, npl, npu, d, al, au, indexl, iteml, indexu, itemu, r, z, iarray, rarray)
    mat%Iarray(97) = 1
    mat%Iarray(98) = 1
    call c_f_pointer(r, rv, [3*np])
    call c_f_pointer(z, zv, [3*np])
    allocate(wk(3*np))
    wk = 0.d0
    tcomm = 0.d0
    call hecmw_precond_setup(mat, mesh, 1)
    call hecmw_precond_apply(mesh, mat, rv, zv, wk, tcomm)
    deallocate(wk)
  end subroutine ref_precond

  !> sqrt(|b - A x|^2 / |b|^2) through hecmw_rel_resid_L2.
  function ref_rel_resid(n, np, npl, npu, d, al, au, indexl, iteml, indexu, itemu, &
      b, x, nthreads) result(res) bind(c, name='ref_rel_resid')
    use hecmw_solver_las
    integer(c_int), value :: n, np, npl, npu, nthreads
    type(c_ptr), value :: d, al, au, indexl, iteml, indexu, itemu, b, x
    real(c_double) :: res
    integer(c_int) :: ia(100)
    real(c_double) :: ra(100)
    ia = 0; ra = 0.d0
    call set_threads(nthreads)
    call bind_matrix(n, np, npl, npu, d, al, au, indexl, iteml, indexu, itemu, b, x, ia, ra)
    res = hecmw_rel_resid_L2(mesh, mat)
  end function ref_rel_resid

  !> hecmw_mat_dump (type 1 MM, 2 CSR, 3 BSR, with the right-hand side) and
  !> hecmw_mat_dump_solution: the files dump_matrix_<call>_0.{mm,csr,bsr,rhs,sol}
  !> land in the current directory (hecmw1/src/solver/matrix/hecmw_matrix_dump.f90).
  subroutine ref_mat_dump(n, np, npl, npu, d, al, au, indexl, iteml, indexu, itemu, &
      b, x, dump_type) bind(c, name='ref_mat_dump')
    use hecmw_matrix_misc
    use hecmw_matrix_dump
    integer(c_int), value :: n, np, npl, npu, dump_type
    type(c_ptr), value :: d, al, au, indexl, iteml, indexu, itemu, b, x
    integer(c_int) :: ia(100)
    real(c_double) :: ra(100)
    ia = 0; ra = 0.d0
    call bind_matrix(n, np, npl, npu, d, al, au, indexl, iteml, indexu, itemu, b, x, ia, ra)
    call hecmw_mat_set_dump(mat, dump_type)
    call hecmw_mat_set_dump_exit(mat, 0)
    call hecmw_mat_dump(mat, mesh)
    call hecmw_mat_dump_solution(mat)
  end subroutine ref_mat_dump

  !> Flush Fortran unit 6 (ITERLOG lines) so a caller capturing stdout sees them.
  subroutine ref_flush() bind(c, name='ref_flush')
    flush(6)
  end subroutine ref_flush

end module hecmw_ref_driver
