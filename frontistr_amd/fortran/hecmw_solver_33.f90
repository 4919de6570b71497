!-------------------------------------------------------------------------------
! hecmw_solve_33: GPU (MI355X) drop-in for hecmw1's NDOF=3 iterative solve.
!
! Same argument list as hecmw_solve_iterative
! (hecmw1/src/solver/iterative/hecmw_solver_Iterative.f90:13) and the name the
! reference's commented-out NDOF dispatch still carries
! (hecmw1/src/solver/main/hecmw_solver.f90:113).  The hecmwST_matrix arrays are
! handed to libhecmw33.so through iso_c_binding without copies; the library
! returns the reference's error codes and writes the same Iarray/Rarray slots
! the reference writes (recycle counters 96-98, converged/diverged 81/82,
! SIGMA_DIAG during the automatic retry).  Here the codes are handled as
! hecmw_solve_iterative handles them: hecmw_solve_error (which stops on
! 1001/2001) and the ITERLOG/TIMELOG lines.
!
! The side effects of hecmw_solve_iterative around the Krylov call are kept:
! hecmw_mat_dump before the solve (DUMPTYPE, Iarray 31; DUMPEXIT stops there),
! hecmw_mat_dump_solution after it (hecmw_solver_Iterative.f90:90,160), and the
! TIMELOG=VERBOSE matvec bandwidth/FLOPs block of hecmw_output_flops (:185,
! :189-254) from the GPU's SpMV time.
!
! hecmw33_available() tells a caller whether the GPU path can take a system
! (library loaded, a device visible, METHOD 1/2, PRECOND 1/2/3, one domain,
! multipoint constraints the MPC-CG product supports; SCALING is applied on
! the device copies); otherwise the caller keeps the reference path.
!-------------------------------------------------------------------------------
module hecmw_solver_33
  use iso_c_binding
  use hecmw_util
  implicit none
  private
  public :: hecmw_solve_33
  public :: hecmw33_available

  type, bind(c) :: hecmw33_sys
    integer(c_int32_t) :: N, NP, NPL, NPU
    type(c_ptr) :: D, AL, AU, indexL, indexU, itemL, itemU
  end type hecmw33_sys

  type, bind(c) :: hecmw33_out
    integer(c_int32_t) :: iter, error, method, ncolor
    real(c_double) :: resid, resid_true, Tset, Tsol, Tcomm, time_Ax, time_precond, t_upload
  end type hecmw33_out

  interface
    function c_hecmw33_solve(A, B, X, iarray, rarray, out) bind(c, name='hecmw33_solve') result(rc)
      import :: hecmw33_sys, hecmw33_out, c_ptr, c_int32_t, c_double
      type(hecmw33_sys), intent(in) :: A
      type(c_ptr), value :: B, X
      integer(c_int32_t), intent(inout) :: iarray(100)
      real(c_double), intent(inout) :: rarray(100)
      type(hecmw33_out), intent(out) :: out
      integer(c_int32_t) :: rc
    end function c_hecmw33_solve

    function c_hecmw33_device_count() bind(c, name='hecmw33_device_count') result(n)
      import :: c_int32_t
      integer(c_int32_t) :: n
    end function c_hecmw33_device_count

    function c_hecmw33_set_mpc(N, n_mpc, idx, item, dof, val) bind(c, name='hecmw33_set_mpc') result(rc)
      import :: c_ptr, c_int32_t
      integer(c_int32_t), value :: N, n_mpc
      type(c_ptr), value :: idx, item, dof, val
      integer(c_int32_t) :: rc
    end function c_hecmw33_set_mpc

    function c_hecmw33_set_coords(NP, xyz) bind(c, name='hecmw33_set_coords') result(rc)
      import :: c_ptr, c_int32_t
      integer(c_int32_t), value :: NP
      type(c_ptr), value :: xyz
      integer(c_int32_t) :: rc
    end function c_hecmw33_set_coords

    function c_hecmw33_last_history(hist, cap) bind(c, name='hecmw33_last_history') result(n)
      import :: c_double, c_int32_t
      real(c_double), intent(out) :: hist(*)
      integer(c_int32_t), value :: cap
      integer(c_int32_t) :: n
    end function c_hecmw33_last_history
  end interface

contains

  logical function hecmw33_available(hecMESH, hecMAT)
    type(hecmwST_local_mesh) :: hecMESH
    type(hecmwST_matrix) :: hecMAT
    integer(kind=kint) :: method, precond
    method = hecMAT%Iarray(2)
    precond = hecMAT%Iarray(3)
    hecmw33_available = hecMAT%NDOF == 3 .and. (method == 1 .or. method == 2) &
      .and. (precond == 1 .or. precond == 2 .or. precond == 3 .or. hecMAT%Iarray(5) <= 0) &
      .and. hecMESH%n_neighbor_pe == 0
    if (hecmw33_available) hecmw33_available = c_hecmw33_device_count() > 0
    ! multipoint constraints: the library takes them for the MPC-CG product
    ! unless they are chained or touch external nodes
    if (hecmw33_available) hecmw33_available = pass_mpc(hecMESH, hecMAT) == 0
  end function hecmw33_available

  ! hecMESH%mpc to the library (n_mpc = 0 clears what an earlier call left)
  integer(c_int32_t) function pass_mpc(hecMESH, hecMAT)
    type(hecmwST_local_mesh), target :: hecMESH
    type(hecmwST_matrix) :: hecMAT
    if (hecMESH%mpc%n_mpc > 0) then
      pass_mpc = c_hecmw33_set_mpc(int(hecMAT%N, c_int32_t), int(hecMESH%mpc%n_mpc, c_int32_t), &
        c_loc(hecMESH%mpc%mpc_index(0)), c_loc(hecMESH%mpc%mpc_item(1)), c_loc(hecMESH%mpc%mpc_dof(1)), &
        c_loc(hecMESH%mpc%mpc_val(1)))
    else
      pass_mpc = c_hecmw33_set_mpc(int(hecMAT%N, c_int32_t), 0_c_int32_t, c_null_ptr, c_null_ptr, &
        c_null_ptr, c_null_ptr)
    endif
  end function pass_mpc

  subroutine hecmw_solve_33(hecMESH, hecMAT)
    use m_hecmw_solve_error
    use hecmw_matrix_dump
    type(hecmwST_local_mesh) :: hecMESH
    type(hecmwST_matrix), target :: hecMAT
    type(hecmw33_sys) :: A
    type(hecmw33_out) :: out
    integer(c_int32_t) :: rc, ia(100), nh, i
    real(c_double) :: ra(100)
    real(c_double), allocatable :: hist(:)
    real(kind=kreal) :: TR

    A%N = hecMAT%N
    A%NP = hecMAT%NP
    A%NPL = hecMAT%indexL(hecMAT%NP)
    A%NPU = hecMAT%indexU(hecMAT%NP)
    A%D = c_loc(hecMAT%D(1))
    A%AL = c_null_ptr
    A%AU = c_null_ptr
    A%itemL = c_null_ptr
    A%itemU = c_null_ptr
    if (A%NPL > 0) then
      A%AL = c_loc(hecMAT%AL(1))
      A%itemL = c_loc(hecMAT%itemL(1))
    endif
    if (A%NPU > 0) then
      A%AU = c_loc(hecMAT%AU(1))
      A%itemU = c_loc(hecMAT%itemU(1))
    endif
    A%indexL = c_loc(hecMAT%indexL(0))
    A%indexU = c_loc(hecMAT%indexU(0))
    ia(:) = hecMAT%Iarray(:)
    ra(:) = hecMAT%Rarray(:)
    rc = pass_mpc(hecMESH, hecMAT)
    if (rc < 0) then
      write(*,'(a,i0)') '#### hecmw_solve_33: multipoint constraints not supported by the GPU library ', rc
      call hecmw_abort(hecmw_comm_get_comm())
    endif

    ! the two-level correction of the flexible CG (SOLVER_OPT(3) = Iarray(43) > 0)
    ! builds its aggregates from the node coordinates (Fortran does not short-circuit
    ! .and., so the pointer is inspected only when the option is on)
    if (ia(43) > 0) then
      if (associated(hecMESH%node)) then
        if (size(hecMESH%node) >= 3 * hecMAT%NP) then
          rc = c_hecmw33_set_coords(int(hecMAT%NP, c_int32_t), c_loc(hecMESH%node(1)))
        endif
      endif
    endif

    ! the system as handed over (DUMPTYPE; the dump numbering is the reference's
    ! own counter, shared with every other hecmw_mat_dump call of the process)
    call hecmw_mat_dump(hecMAT, hecMESH)

    rc = c_hecmw33_solve(A, c_loc(hecMAT%B(1)), c_loc(hecMAT%X(1)), ia, ra, out)
    hecMAT%Iarray(:) = ia(:)
    hecMAT%Rarray(:) = ra(:)

    if (rc < 0) then
      write(*,'(a,i0)') '#### hecmw_solve_33: GPU library failure ', rc
      call hecmw_abort(hecmw_comm_get_comm())
    endif
    if (rc == HECMW_SOLVER_ERROR_ZERO_DIAG .or. rc == HECMW_SOLVER_ERROR_INCONS_PC) then
      call hecmw_solve_error(hecMESH, rc)
    endif

    ! per-iteration residuals, as hecmw_solve_CG prints them (hecmw_solver_CG.f90:253)
    if (hecMESH%my_rank == 0 .and. hecMAT%Iarray(21) == 1) then
      allocate(hist(max(1, out%iter)))
      nh = c_hecmw33_last_history(hist, int(size(hist), c_int32_t))
      do i = 1, nh
        write(*,'(i7, 1pe16.6)') i, hist(i)
      enddo
      deallocate(hist)
    endif

    if (out%error /= 0) call hecmw_solve_error(hecMESH, out%error)

    ! post-solve lines of hecmw_solve_iterative (hecmw_solver_Iterative.f90:155-183)
    if (hecMESH%my_rank == 0 .and. (hecMAT%Iarray(21) == 1 .or. hecMAT%Iarray(22) >= 1)) then
      write(*,"(a,1pe12.5)") '### Relative residual =', out%resid_true
    endif
    call hecmw_mat_dump_solution(hecMAT)
    if (hecMESH%my_rank == 0 .and. hecMAT%Iarray(22) >= 1) then
      TR = (out%Tsol - out%Tcomm) / (out%Tsol + 1.d-24) * 100.d0
      write(*,'(/a)')          '### summary of linear solver'
      write(*,'(i10,a, 1pe16.6)')      out%iter, ' iterations  ', out%resid
      write(*,'(a, 1pe16.6 )') '    set-up time      : ', out%Tset
      write(*,'(a, 1pe16.6 )') '    solver time      : ', out%Tsol
      write(*,'(a, 1pe16.6 )') '    solver/comm time : ', out%Tcomm
      write(*,'(a, 1pe16.6 )') '    solver/matvec    : ', out%time_Ax
      write(*,'(a, 1pe16.6 )') '    solver/precond   : ', out%time_precond
      if (out%iter > 0) write(*,'(a, 1pe16.6 )') '    solver/1 iter    : ', out%Tsol / out%iter
      write(*,'(a, 1pe16.6/)') '    work ratio (%)   : ', TR
    endif
    if (hecMAT%Iarray(22) == 2) call output_flops(hecMESH, hecMAT, out%iter, out%time_Ax)
  end subroutine hecmw_solve_33

  ! TIMELOG=VERBOSE: the matrix's memory amount, the matvec bandwidth and FLOP
  ! rate over count_Ax products taking time_Ax seconds, with their statistics
  ! over the ranks, in hecmw_output_flops' lines and units
  ! (hecmw_solver_Iterative.f90:189-254: bytes 8*9*NZ + 4*(NPU+NPL) + 8*(N+1)
  ! + 24*N + 24*NP, 18*NZ flops, NZ = N + NPU + NPL, SI prefixes in 1000s).
  subroutine output_flops(hecMESH, hecMAT, count_Ax, time_Ax)
    use m_hecmw_comm_f
    use hecmw_solver_misc
    type(hecmwST_local_mesh) :: hecMESH
    type(hecmwST_matrix) :: hecMAT
    integer(c_int32_t), intent(in) :: count_Ax
    real(c_double), intent(in) :: time_Ax
    character(2), parameter :: prefix(0:6) = ['  ', ' K', ' M', ' G', ' T', ' P', ' E']
    character(len=29), parameter :: label(2) = ['matvec memory band width     ', 'matvec FLOPs                 ']
    character(len=5), parameter :: unit(2) = ['B/s  ', 'FLOPs']
    integer(kind=kint) :: NZ, k, e
    real(kind=kreal) :: bytes, flops, rate, scaled, t_max, t_min, t_avg, t_sd

    NZ = hecMAT%N + hecMAT%indexU(hecMAT%N) + hecMAT%indexL(hecMAT%N)
    bytes = dble(kreal) * NZ * 9 + dble(kint) * (hecMAT%indexU(hecMAT%N) + hecMAT%indexL(hecMAT%N)) &
          + dble(kint) * 2 * (hecMAT%N + 1) + dble(kreal) * 3 * (hecMAT%N + hecMAT%NP)
    flops = 2.0d0 * NZ * 9
    call hecmw_allreduce_R1(hecMESH, bytes, hecmw_sum)
    call hecmw_allreduce_R1(hecMESH, flops, hecmw_sum)
    call si_scale(bytes, scaled, e)
    if (hecMESH%my_rank == 0) write(*, "(a,f11.3,a,a)") 'memory amount of coef. matrix: ', scaled, prefix(e), 'B'
    do k = 1, 2
      if (k == 1) then
        rate = count_Ax * bytes / time_Ax
      else
        rate = count_Ax * flops / time_Ax
      endif
      call si_scale(rate, scaled, e)
      call hecmw_time_statistics(hecMESH, scaled, t_max, t_min, t_avg, t_sd)
      if (hecMESH%my_rank == 0) then
        write(*, "(a,f11.3,a,a)") label(k) // ': ', scaled, prefix(e), trim(unit(k))
        write(*, "(a,f11.3)") '  Max     :', t_max
        write(*, "(a,f11.3)") '  Min     :', t_min
        write(*, "(a,f11.3)") '  Avg     :', t_avg
        write(*, "(a,f11.3)") '  Std Dev :', t_sd
      endif
    enddo
  contains
    subroutine si_scale(v, scaled, e)
      real(kind=kreal), intent(in) :: v
      real(kind=kreal), intent(out) :: scaled
      integer(kind=kint), intent(out) :: e
      e = int(log(v) / log(1000.d0))
      scaled = v / 1000.d0**e
    end subroutine si_scale
  end subroutine output_flops

end module hecmw_solver_33
