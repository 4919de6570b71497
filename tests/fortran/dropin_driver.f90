!-------------------------------------------------------------------------------
! Drop-in check (test infrastructure): one hecmwST_matrix, solved once by the
! reference's own hecmw_solve_iterative (CPU, compiled from the FrontISTR
! sources) and once by hecmw_solve_33 (GPU, frontistr_amd/fortran), both
! called from Fortran exactly as fistr1 would call them.
!
!   dropin_test <n> <method> <precond> [scaling] [fcg] [mpc] [extra]    (OMP_NUM_THREADS >= 2)
! mpc = 1: the z displacement of every top-face node (but one corner) is tied
! to the top-face centre node's (hecMESH%mpc, MPC_METHOD = 2), B replaced by
! T^T B as fistr1 hands it over; both sides then solve with T^T A T products
! fcg = 1: the GPU side runs the flexible CG with FP32 operators and the
! two-level correction (Iarray 41 = 1, 43 = 1), with hecMESH%node set
! extra = 1: both sides with DUMPTYPE=BSR and TIMELOG=VERBOSE, so each writes
! dump_matrix_<call>_0.{bsr,rhs,sol} into the working directory (call 1 the
! reference's, 2 the GPU's) and prints its summary and matvec bandwidth lines
! extra = 2: a Newton-like sequence of six solves of the same profile with new
! values each time (stiffer, plus a diagonal term), Iarray(97) = 2 on the first
! call and 1 afterwards (fstr_solve_NonLinear.f90:127-131), MAXRECYCLE = 3;
! one line per call:  DROPIN_SEQ k nrecycle numfact symbfact (reference, GPU) relerr
! prints one line:  DROPIN n method precond conv_ref conv_gpu div_ref div_gpu relerr
!-------------------------------------------------------------------------------
program dropin_test
  use iso_c_binding
  use hecmw_util
  use hecmw_matrix_misc
  use hecmw_solver_iterative
  use hecmw_solver_33
  implicit none

  interface
    function gen_count(n, N_, NPL, NPU) bind(c, name='hecmw33_gen_hex8_count') result(rc)
      import :: c_int32_t
      integer(c_int32_t), value :: n
      integer(c_int32_t), intent(out) :: N_, NPL, NPU
      integer(c_int32_t) :: rc
    end function gen_count
    function gen(n, kind, p, seed, D, AL, AU, iL, tL, iU, tU, B, E) bind(c, name='hecmw33_gen_hex8') result(rc)
      import :: c_int32_t, c_double, c_int64_t, c_ptr
      integer(c_int32_t), value :: n, kind
      real(c_double), intent(in) :: p(*)
      integer(c_int64_t), value :: seed
      real(c_double), intent(out) :: D(*), AL(*), AU(*), B(*)
      integer(c_int32_t), intent(out) :: iL(*), tL(*), iU(*), tU(*)
      type(c_ptr), value :: E
      integer(c_int32_t) :: rc
    end function gen
  end interface

  type(hecmwST_matrix), target :: mref, mgpu
  type(hecmwST_local_mesh) :: mesh
  integer(c_int32_t) :: n, N_, NPL, NPU, rc
  integer :: method, precond, scaling, fcg, mpc, extra, i, j, k, nn, c, s, t, ks, km, call_k
  real(kind=kreal), allocatable :: D0(:), AL0(:), AU0(:), B0(:)
  real(kind=kreal) :: f, dmean
  character(len=32) :: arg
  real(c_double) :: p(8)
  real(kind=kreal) :: relerr

  call get_command_argument(1, arg); read(arg, *) n
  call get_command_argument(2, arg); read(arg, *) method
  call get_command_argument(3, arg); read(arg, *) precond
  scaling = 0
  fcg = 0
  if (command_argument_count() >= 4) then
    call get_command_argument(4, arg); read(arg, *) scaling
  endif
  if (command_argument_count() >= 5) then
    call get_command_argument(5, arg); read(arg, *) fcg
  endif
  mpc = 0
  if (command_argument_count() >= 6) then
    call get_command_argument(6, arg); read(arg, *) mpc
  endif
  extra = 0
  if (command_argument_count() >= 7) then
    call get_command_argument(7, arg); read(arg, *) extra
  endif

  rc = gen_count(n, N_, NPL, NPU)
  call setup(mref)
  call setup(mgpu)
  p = 0.d0
  p(1) = 210000.d0
  p(2) = 0.3d0
  rc = gen(n, 0_c_int32_t, p, 0_c_int64_t, mref%D, mref%AL, mref%AU, mref%indexL, mref%itemL, &
           mref%indexU, mref%itemU, mref%B, c_null_ptr)
  mgpu%D = mref%D
  mgpu%AL = mref%AL
  mgpu%AU = mref%AU
  mgpu%indexL = mref%indexL
  mgpu%indexU = mref%indexU
  mgpu%itemL = mref%itemL
  mgpu%itemU = mref%itemU
  mgpu%B = mref%B

  mesh%nn_internal = N_
  mesh%n_node = N_
  mesh%n_dof = 3
  mesh%n_neighbor_pe = 0
  mesh%my_rank = 0
  mesh%PETOT = 1
  mesh%zero = 1
  mesh%mpc%n_mpc = 0
  nullify(mesh%node)
  if (fcg == 1) then
    ! node i + (n+1)(j + (n+1)k) of the generated cube sits at (i, j, k)/n
    nn = n + 1
    allocate(mesh%node(3*N_))
    do k = 0, n
      do j = 0, n
        do i = 0, n
          mesh%node(3*(i + nn*(j + nn*k)) + 1) = dble(i) / n
          mesh%node(3*(i + nn*(j + nn*k)) + 2) = dble(j) / n
          mesh%node(3*(i + nn*(j + nn*k)) + 3) = dble(k) / n
        enddo
      enddo
    enddo
    mgpu%Iarray(41) = 1
    mgpu%Iarray(43) = 1
  endif
  mesh%mpc%n_mpc = 0
  if (mpc == 1) then
    nn = n + 1
    c = n / 2
    t = 0
    allocate(mesh%mpc%mpc_index(0:nn*nn), mesh%mpc%mpc_item(2*nn*nn), mesh%mpc%mpc_dof(2*nn*nn), &
             mesh%mpc%mpc_val(2*nn*nn))
    mesh%mpc%mpc_index(0) = 0
    km = 1 + c + nn * (c + nn * n)
    do j = 0, n
      do i = 0, n
        s = 1 + i + nn * (j + nn * n)
        if (s == km .or. (i == 0 .and. j == 0)) cycle
        mesh%mpc%mpc_item(2*t+1) = s
        mesh%mpc%mpc_item(2*t+2) = km
        mesh%mpc%mpc_dof(2*t+1:2*t+2) = 3
        mesh%mpc%mpc_val(2*t+1) = 1.d0
        mesh%mpc%mpc_val(2*t+2) = -1.d0
        t = t + 1
        mesh%mpc%mpc_index(t) = 2 * t
        ! B <- T^T B: the slave's load moves to the master
        ks = 3 * (s - 1) + 3
        mref%B(3 * (km - 1) + 3) = mref%B(3 * (km - 1) + 3) + mref%B(ks)
        mref%B(ks) = 0.d0
      enddo
    enddo
    mesh%mpc%n_mpc = t
    mgpu%B = mref%B
    call hecmw_mat_set_mpc_method(mref, 2)
    call hecmw_mat_set_mpc_method(mgpu, 2)
  endif

  if (extra == 1) then
    mref%Iarray(31) = 3
    mgpu%Iarray(31) = 3
    call hecmw_mat_set_timelog(mref, 2)
    call hecmw_mat_set_timelog(mgpu, 2)
  endif

  if (extra == 2) then
    allocate(D0(size(mref%D)), AL0(size(mref%AL)), AU0(size(mref%AU)), B0(size(mref%B)))
    D0 = mref%D
    AL0 = mref%AL
    AU0 = mref%AU
    B0 = mref%B
    dmean = sum(D0(1::9)) / N_
    relerr = 0.d0
    do call_k = 0, 5
      f = 1.d0 + 0.2d0 * call_k
      mref%AL = AL0 * f
      mref%AU = AU0 * f
      mref%D = D0
      do i = 1, N_
        if (D0(9*i-8) /= 1.d0) then   ! a free node (fixed ones keep their unit rows)
          mref%D(9*i-8:9*i) = D0(9*i-8:9*i) * f
          mref%D(9*i-8) = mref%D(9*i-8) + 0.05d0 * call_k * dmean
          mref%D(9*i-4) = mref%D(9*i-4) + 0.05d0 * call_k * dmean
          mref%D(9*i) = mref%D(9*i) + 0.05d0 * call_k * dmean
        endif
      enddo
      mref%B = B0 * (1.d0 + 0.1d0 * call_k)
      mgpu%D = mref%D
      mgpu%AL = mref%AL
      mgpu%AU = mref%AU
      mgpu%B = mref%B
      do k = 1, 2
        if (k == 1) then
          call newton_flags(mref, call_k)
        else
          call newton_flags(mgpu, call_k)
        endif
      enddo
      mref%X = 0.d0
      mgpu%X = 0.d0
      call hecmw_solve_iterative(mesh, mref)
      if (.not. hecmw33_available(mesh, mgpu)) then
        write(*,'(a)') 'DROPIN unavailable'
        stop 1
      endif
      call hecmw_solve_33(mesh, mgpu)
      write(*,'(a,1x,i0,6(1x,i0),1x,es12.4)') 'DROPIN_SEQ', call_k, mref%Iarray(96:98), mgpu%Iarray(96:98), &
        sqrt(sum((mgpu%X - mref%X)**2) / sum(mref%X**2))
      relerr = max(relerr, sqrt(sum((mgpu%X - mref%X)**2) / sum(mref%X**2)))
    enddo
    write(*,'(a,3(1x,i0),4(1x,i0),1x,es12.4)') 'DROPIN', n, method, precond, &
      hecmw_mat_get_flag_converged(mref), hecmw_mat_get_flag_converged(mgpu), &
      hecmw_mat_get_flag_diverged(mref), hecmw_mat_get_flag_diverged(mgpu), relerr
    stop
  endif

  call hecmw_solve_iterative(mesh, mref)
  if (.not. hecmw33_available(mesh, mgpu)) then
    write(*,'(a)') 'DROPIN unavailable'
    stop 1
  endif
  call hecmw_solve_33(mesh, mgpu)
  relerr = sqrt(sum((mgpu%X - mref%X)**2) / sum(mref%X**2))
  write(*,'(a,3(1x,i0),4(1x,i0),1x,es12.4)') 'DROPIN', n, method, precond, &
    hecmw_mat_get_flag_converged(mref), hecmw_mat_get_flag_converged(mgpu), &
    hecmw_mat_get_flag_diverged(mref), hecmw_mat_get_flag_diverged(mgpu), relerr

contains

  ! the recycle flags fistr1's Newton loop sets before each solve
  subroutine newton_flags(m, it)
    type(hecmwST_matrix) :: m
    integer :: it
    if (it == 0) then
      m%Iarray(98) = 1
      m%Iarray(97) = 2
    else
      m%Iarray(97) = 1
    endif
    m%Iarray(35) = 3
  end subroutine newton_flags

  subroutine setup(m)
    type(hecmwST_matrix), target :: m
    call hecmw_mat_init(m)
    m%N = N_
    m%NP = N_
    m%NPL = NPL
    m%NPU = NPU
    m%NDOF = 3
    allocate(m%D(9*N_), m%AL(9*NPL), m%AU(9*NPU), m%B(3*N_), m%X(3*N_))
    allocate(m%indexL(0:N_), m%indexU(0:N_), m%itemL(NPL), m%itemU(NPU))
    m%X = 0.d0
    call hecmw_mat_set_iter(m, 5000)
    call hecmw_mat_set_method(m, method)
    call hecmw_mat_set_precond(m, precond)
    call hecmw_mat_set_iterlog(m, 0)
    call hecmw_mat_set_timelog(m, 0)
    call hecmw_mat_set_scaling(m, scaling)
  end subroutine setup

end program dropin_test
