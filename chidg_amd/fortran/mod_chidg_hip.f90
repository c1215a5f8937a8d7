!> Fortran binding of the gfx950 linear-solve library (include/chidg_hip.h).
!!
!! This is the iso_c_binding layer a ChiDG build with backend 'hip' uses in
!! place of the native containers on the Newton-Krylov linear solve:
!!   hip_matrix_t   <- chidg_matrix_t   (src/containers/type_chidg_matrix.f90:38-83)
!!   hip_mv         <- chidg_mv         (src/containers/operator_chidg_mv.f90:46-61)
!!   hip_precon_t   <- precon_ILU0_t / precon_identity_t
!!                     init/update/apply/tear_down (src/preconditioners/precon_ILU0.f90:64-470)
!!   hip_fgmres_t   <- fgmres_t with solve(A,x,b,M) (src/linear_solvers/type_fgmres.f90:81-310)
!!   hip_vector_t   <- chidg_vector_t with backend 'hip': values resident on the
!!                     device between calls, with the vector operations a
!!                     Newton loop needs (src/containers/type_chidg_vector.f90:
!!                     norm :1237-1300, the overloaded +, -, * :1740-2080)
!! Blocks are the reference's densematrix_t%mat: n x n column-major real(8);
!! vectors are contiguous element-major real(8) arrays.  Every call aborts
!! through hip_check on a non-zero status, the way chidg_signal(FATAL, ...)
!! ends the reference (src/tools/messenger.f90:295-304).
module mod_chidg_hip
    use iso_c_binding
    implicit none
    private

    integer(c_int32_t), parameter, public :: HIP_KIND_LOCAL = 0, HIP_KIND_COUPLING = 1
    integer(c_int32_t), parameter, public :: HIP_PREC_IDENTITY = 0, HIP_PREC_ILU0 = 1, HIP_PREC_JACOBI = 2
    integer(c_int32_t), parameter, public :: HIP_FP64 = 0, HIP_FP32 = 1, HIP_BF16 = 2
    integer(c_int32_t), parameter, public :: HIP_CGS = 0, HIP_MGS = 1
    ! ILU(0) apply schedules (cdg_prec_set_schedule); AUTO is the default
    integer(c_int32_t), parameter, public :: HIP_SCHEDULE_LEVELS = 0, HIP_SCHEDULE_DATAFLOW = 1, &
                                             HIP_SCHEDULE_FOLDED = 2, HIP_SCHEDULE_AUTO = 3, &
                                             HIP_SCHEDULE_STREAM = 4, HIP_SCHEDULE_ROWSLICE = 5

    ! cdg_fgmres_params / cdg_fgmres_stats (same layout as the C structs)
    type, bind(c), public :: hip_fgmres_params_t
        integer(c_int32_t) :: nkrylov
        real(c_double)     :: tol
        real(c_double)     :: rtol
        integer(c_int32_t) :: nmax
        integer(c_int32_t) :: orthogonalization
        integer(c_int32_t) :: inner_fgmres
        integer(c_int32_t) :: inner_nkrylov
        real(c_double)     :: inner_tol
        real(c_double)     :: inner_rtol
        integer(c_int32_t) :: inner_nmax
        integer(c_int32_t) :: inner_orthogonalization
        integer(c_int32_t) :: update_preconditioner
    end type hip_fgmres_params_t

    type, bind(c), public :: hip_fgmres_stats_t
        integer(c_int32_t) :: niter
        real(c_double)     :: residual_norm
        integer(c_int32_t) :: nrestart
        integer(c_int32_t) :: nreorth
        integer(c_int32_t) :: factored
        real(c_double)     :: t_total, t_factor, t_precon, t_mv, t_dot, t_norm
    end type hip_fgmres_stats_t

    interface
        integer(c_int) function cdg_ctx_create(device, ctx) bind(c, name='cdg_ctx_create')
            import :: c_int, c_int32_t, c_ptr
            integer(c_int32_t), value :: device
            type(c_ptr) :: ctx
        end function
        integer(c_int) function cdg_ctx_destroy(ctx) bind(c, name='cdg_ctx_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        type(c_ptr) function cdg_last_error(ctx) bind(c, name='cdg_last_error')
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function cdg_mat_create(ctx, nrows, n, row_ptr, col, kind, mat) bind(c, name='cdg_mat_create')
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nrows, n
            integer(c_int32_t), intent(in) :: row_ptr(*), col(*), kind(*)
            type(c_ptr) :: mat
        end function
        integer(c_int) function cdg_mat_destroy(ctx, mat) bind(c, name='cdg_mat_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, mat
        end function
        integer(c_int) function cdg_mat_upload(ctx, mat, blocks) bind(c, name='cdg_mat_upload')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, mat
            real(c_double), intent(in) :: blocks(*)
        end function
        integer(c_int) function cdg_mat_download(ctx, mat, blocks) bind(c, name='cdg_mat_download')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, mat
            real(c_double), intent(inout) :: blocks(*)
        end function
        integer(c_int) function cdg_mat_tune(ctx, mat, reps, best_ms) bind(c, name='cdg_mat_tune')
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx, mat
            integer(c_int32_t), value :: reps
            real(c_double) :: best_ms
        end function
        integer(c_int) function cdg_ctx_mem_used(ctx, bytes) bind(c, name='cdg_ctx_mem_used')
            import :: c_int, c_int64_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int64_t) :: bytes
        end function
        integer(c_int) function cdg_mv(ctx, mat, x, y) bind(c, name='cdg_mv')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, mat
            real(c_double), intent(in) :: x(*)
            real(c_double), intent(inout) :: y(*)
        end function
        integer(c_int) function cdg_prec_create(ctx, mat, ptype, precision, prec) bind(c, name='cdg_prec_create')
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx, mat
            integer(c_int32_t), value :: ptype, precision
            type(c_ptr) :: prec
        end function
        integer(c_int) function cdg_prec_update(ctx, prec, mat, bad_row) bind(c, name='cdg_prec_update')
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx, prec, mat
            integer(c_int32_t) :: bad_row
        end function
        integer(c_int) function cdg_prec_apply(ctx, prec, mat, v, z) bind(c, name='cdg_prec_apply')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, prec, mat
            real(c_double), intent(in) :: v(*)
            real(c_double), intent(inout) :: z(*)
        end function
        integer(c_int) function cdg_prec_set_schedule(ctx, prec, schedule) bind(c, name='cdg_prec_set_schedule')
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx, prec
            integer(c_int32_t), value :: schedule
        end function
        integer(c_int) function cdg_prec_tune(ctx, prec, mat, reps, best_ms) bind(c, name='cdg_prec_tune')
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx, prec, mat
            integer(c_int32_t), value :: reps
            real(c_double) :: best_ms
        end function
        integer(c_int) function cdg_prec_destroy(ctx, prec) bind(c, name='cdg_prec_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, prec
        end function
        integer(c_int) function cdg_prec_download(ctx, prec, L, Dinv) bind(c, name='cdg_prec_download')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, prec
            real(c_double), intent(inout) :: L(*), Dinv(*)
        end function
        integer(c_int) function cdg_fgmres_solve(ctx, mat, prec, b, x, params, stats) bind(c, name='cdg_fgmres_solve')
            import :: c_int, c_ptr, c_double, hip_fgmres_params_t, hip_fgmres_stats_t
            type(c_ptr), value :: ctx, mat, prec
            real(c_double), intent(in) :: b(*)
            real(c_double), intent(inout) :: x(*)
            type(hip_fgmres_params_t), intent(in) :: params
            type(hip_fgmres_stats_t), intent(inout) :: stats
        end function
        subroutine cdg_fgmres_default_params(params) bind(c, name='cdg_fgmres_default_params')
            import :: hip_fgmres_params_t
            type(hip_fgmres_params_t), intent(inout) :: params
        end subroutine
        integer(c_int) function cdg_mat_stamp(mat, stamp) bind(c, name='cdg_mat_stamp')
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value :: mat
            integer(c_int64_t) :: stamp
        end function
        integer(c_int) function cdg_vec_create(ctx, len, v) bind(c, name='cdg_vec_create')
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: len
            type(c_ptr) :: v
        end function
        integer(c_int) function cdg_vec_destroy(ctx, v) bind(c, name='cdg_vec_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, v
        end function
        integer(c_int) function cdg_vec_copy(ctx, len, src, dst) bind(c, name='cdg_vec_copy')
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value :: ctx, src, dst
            integer(c_int64_t), value :: len
        end function
        integer(c_int) function cdg_vec_axpby(ctx, len, a, x, b, y) bind(c, name='cdg_vec_axpby')
            import :: c_int, c_ptr, c_int64_t, c_double
            type(c_ptr), value :: ctx, x, y
            integer(c_int64_t), value :: len
            real(c_double), value :: a, b
        end function
        integer(c_int) function cdg_vec_dot(ctx, len, x, y, out) bind(c, name='cdg_vec_dot')
            import :: c_int, c_ptr, c_int64_t, c_double
            type(c_ptr), value :: ctx, x, y
            integer(c_int64_t), value :: len
            real(c_double) :: out
        end function
        integer(c_int) function cdg_vec_norm(ctx, len, x, out) bind(c, name='cdg_vec_norm')
            import :: c_int, c_ptr, c_int64_t, c_double
            type(c_ptr), value :: ctx, x
            integer(c_int64_t), value :: len
            real(c_double) :: out
        end function
        integer(c_int) function cdg_euler_residual_gpu(ctx, nx, ny, nz, order, q, R) &
                bind(c, name='cdg_euler_residual_gpu')
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx, q, R
            integer(c_int32_t), value :: nx, ny, nz, order
        end function
        integer(c_size_t) function c_strlen(s) bind(c, name='strlen')
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function
    end interface

    !> One device + stream (cdg_ctx).
    type, public :: hip_context_t
        type(c_ptr) :: h = c_null_ptr
    contains
        procedure :: init => ctx_init
        procedure :: mem_used => ctx_mem_used
        procedure :: release => ctx_release
    end type

    !> Block-sparse matrix resident on the device (chidg_matrix_t, backend 'hip').
    type, public :: hip_matrix_t
        type(c_ptr) :: h = c_null_ptr
        type(c_ptr) :: ctx = c_null_ptr
        integer(c_int32_t) :: nrows = 0, n = 0, nnzb = 0
    contains
        procedure :: init => mat_init
        procedure :: store => mat_store       ! upload + stamp: chidg_matrix%assemble
        procedure :: fetch => mat_fetch       ! device blocks back to the host
        procedure :: tune => mat_tune         ! matvec launch-shape autotuner
        procedure :: stamp => mat_stamp       ! chidg_matrix_t%stamp: changes with every store
        procedure :: release => mat_release
    end type

    !> One block row in the layout of domain_matrix_t (type_domain_matrix.f90:32-79):
    !! the lblks(ielem,itime) entries -- densematrix_vector_t%data_(:), each with
    !! its local parent element eparent_l and its n x n matrix mat -- followed by
    !! the coupling-only entries (chi_blks / bc_blks).  Parents are 1-based
    !! element indices, as in the reference.
    type, public :: hip_block_row_t
        integer(c_int32_t), allocatable :: lparent(:)     ! lblks: eparent_l of data_(k)
        real(c_double),     allocatable :: lmat(:,:,:)    ! lblks: data_(k)%mat = lmat(:,:,k)
        integer(c_int32_t), allocatable :: cparent(:)     ! coupling entries: parent elements
        real(c_double),     allocatable :: cmat(:,:,:)    ! coupling entries: matrices
    end type

    !> One element vector (densevector_t%vec, type_densevector.f90:21-69).
    type, public :: hip_densevector_t
        real(c_double), allocatable :: vec(:)
    end type

    !> Preconditioner (precon_ILU0_t / precon_identity_t).
    type, public :: hip_precon_t
        type(c_ptr) :: h = c_null_ptr
        type(c_ptr) :: ctx = c_null_ptr
        integer(c_int32_t) :: ptype = HIP_PREC_ILU0
        integer(c_int32_t) :: precision = HIP_FP64
        logical :: initialized = .false.
        integer(c_int64_t) :: stamp = 0       ! stamp of the matrix the factor was computed from
    contains
        procedure :: init => prec_init
        procedure :: update => prec_update
        procedure :: apply => prec_apply
        procedure :: set_schedule => prec_set_schedule
        procedure :: tune => prec_tune        ! apply launch-shape autotuner (after update)
        procedure :: tear_down => prec_tear_down
    end type

    !> solver_controller_t (type_solver_controller.f90:25-124): refactor when the
    !! lhs changed since the last call, when forced, or when the preconditioner
    !! was computed from another version of the lhs.
    type, public :: hip_solver_controller_t
        logical :: force_update_lhs = .false.
        logical :: lhs_updated = .false.
        logical :: force_update_preconditioner = .false.
        integer(c_int64_t) :: lhs_stamp = 0
        logical :: preconditioner_updated = .false.
    contains
        procedure :: update_lhs => ctl_update_lhs
        procedure :: update_preconditioner => ctl_update_preconditioner
    end type

    !> fgmres_t: parameters as in linear_solver_t (type_linear_solver.f90:21-40).
    type, public :: hip_fgmres_t
        type(hip_fgmres_params_t) :: params
        integer(c_int32_t) :: niter = 0
        real(c_double) :: residual_norm = 0.d0
        type(hip_fgmres_stats_t) :: stats
    contains
        procedure :: init => fgmres_init
        procedure :: solve => fgmres_solve
        procedure :: solve_dev => fgmres_solve_dev   ! x, b resident on the device
    end type

    !> chidg_vector_t with backend 'hip': nrows*n values on the device, in the
    !! element-major order of get_packed/set_packed.  set/get move values
    !! between host and device; the operations run on the device.
    type, public :: hip_vector_t
        type(c_ptr) :: dev = c_null_ptr
        type(c_ptr) :: ctx = c_null_ptr
        integer(c_int64_t) :: len = 0
    contains
        procedure :: init => vec_init
        procedure :: set => vec_set           ! host values in
        procedure :: get => vec_get           ! device values out
        procedure :: axpby => vec_axpby       ! self = a x + b self
        procedure :: norm => vec_norm         ! chidg_vector_t%norm (one rank)
        procedure :: dot => vec_dot           ! dot(self, x) (operator_chidg_dot.f90:30-89)
        procedure :: release => vec_release
    end type

    public :: hip_mv, hip_check, hip_euler_residual
    public :: hip_pattern_from_rows, hip_pack_blocks, hip_get_packed, hip_set_packed

contains

    subroutine hip_check(status, what, ctx)
        integer(c_int), intent(in) :: status
        character(*), intent(in) :: what
        type(c_ptr), intent(in) :: ctx
        character(len=512) :: msg
        type(c_ptr) :: cmsg
        character(kind=c_char), pointer :: chars(:)
        integer :: i, nchar
        if (status == 0) return
        msg = ''
        if (c_associated(ctx)) then
            cmsg = cdg_last_error(ctx)
            nchar = int(c_strlen(cmsg))
            call c_f_pointer(cmsg, chars, [max(nchar, 1)])
            do i = 1, min(nchar, len(msg))
                msg(i:i) = chars(i)
            end do
        end if
        write(*, '(a, a, a, i0, a, a)') 'FATAL: ', what, ' failed with status ', status, ': ', trim(msg)
        error stop 1
    end subroutine hip_check

    subroutine ctx_init(self, device)
        class(hip_context_t), intent(inout) :: self
        integer, intent(in), optional :: device
        integer(c_int32_t) :: dev
        dev = 0
        if (present(device)) dev = int(device, c_int32_t)
        call hip_check(cdg_ctx_create(dev, self%h), 'cdg_ctx_create', c_null_ptr)
    end subroutine

    !> Bytes of device memory the context holds (matrices, factors, Krylov bases).
    function ctx_mem_used(self) result(bytes)
        class(hip_context_t), intent(in) :: self
        integer(c_int64_t) :: bytes
        bytes = 0
        call hip_check(cdg_ctx_mem_used(self%h, bytes), 'cdg_ctx_mem_used', self%h)
    end function

    subroutine ctx_release(self)
        class(hip_context_t), intent(inout) :: self
        integer(c_int) :: st
        if (c_associated(self%h)) st = cdg_ctx_destroy(self%h)
        self%h = c_null_ptr
    end subroutine

    !> Pattern in ChiDG row order: per block row the lblks entries then the
    !! coupling-only entries; row_ptr is 0-based (C), col holds 0-based parents.
    subroutine mat_init(self, ctx, nrows, n, row_ptr, col, kind)
        class(hip_matrix_t), intent(inout) :: self
        type(hip_context_t), intent(in) :: ctx
        integer, intent(in) :: nrows, n
        integer(c_int32_t), intent(in) :: row_ptr(:), col(:), kind(:)
        self%ctx = ctx%h
        self%nrows = int(nrows, c_int32_t)
        self%n = int(n, c_int32_t)
        self%nnzb = row_ptr(nrows + 1)
        call hip_check(cdg_mat_create(ctx%h, self%nrows, self%n, row_ptr, col, kind, self%h), 'cdg_mat_create', ctx%h)
    end subroutine

    subroutine mat_store(self, blocks)
        class(hip_matrix_t), intent(inout) :: self
        real(c_double), intent(in) :: blocks(:)
        call hip_check(cdg_mat_upload(self%ctx, self%h, blocks), 'cdg_mat_upload', self%ctx)
    end subroutine

    subroutine mat_fetch(self, blocks)
        class(hip_matrix_t), intent(in) :: self
        real(c_double), intent(inout) :: blocks(:)
        call hip_check(cdg_mat_download(self%ctx, self%h, blocks), 'cdg_mat_download', self%ctx)
    end subroutine

    !> Times matvec launch shapes (reps calls each) and keeps the fastest for this
    !! matrix and later ones of the same block size; returns its ms per call.
    function mat_tune(self, reps) result(best_ms)
        class(hip_matrix_t), intent(inout) :: self
        integer, intent(in) :: reps
        real(c_double) :: best_ms
        best_ms = 0.d0
        call hip_check(cdg_mat_tune(self%ctx, self%h, int(reps, c_int32_t), best_ms), 'cdg_mat_tune', self%ctx)
    end function

    function mat_stamp(self) result(stamp)
        class(hip_matrix_t), intent(in) :: self
        integer(c_int64_t) :: stamp
        stamp = 0
        call hip_check(cdg_mat_stamp(self%h, stamp), 'cdg_mat_stamp', self%ctx)
    end function

    subroutine mat_release(self)
        class(hip_matrix_t), intent(inout) :: self
        integer(c_int) :: st
        if (c_associated(self%h)) st = cdg_mat_destroy(self%ctx, self%h)
        self%h = c_null_ptr
    end subroutine

    !> y = A x (chidg_mv)
    function hip_mv(A, x) result(y)
        type(hip_matrix_t), intent(in) :: A
        real(c_double), intent(in) :: x(:)
        real(c_double), allocatable :: y(:)
        allocate(y(size(x)))
        call hip_check(cdg_mv(A%ctx, A%h, x, y), 'cdg_mv', A%ctx)
    end function

    subroutine prec_init(self, A, ptype, precision)
        class(hip_precon_t), intent(inout) :: self
        type(hip_matrix_t), intent(in) :: A
        integer, intent(in), optional :: ptype, precision
        if (present(ptype)) self%ptype = int(ptype, c_int32_t)
        if (present(precision)) self%precision = int(precision, c_int32_t)
        self%ctx = A%ctx
        call hip_check(cdg_prec_create(A%ctx, A%h, self%ptype, self%precision, self%h), 'cdg_prec_create', A%ctx)
        self%initialized = .true.
    end subroutine

    subroutine prec_update(self, A)
        class(hip_precon_t), intent(inout) :: self
        type(hip_matrix_t), intent(in) :: A
        integer(c_int32_t) :: bad
        call hip_check(cdg_prec_update(A%ctx, self%h, A%h, bad), 'cdg_prec_update', A%ctx)
        self%stamp = A%stamp()
    end subroutine

    function prec_apply(self, A, v) result(z)
        class(hip_precon_t), intent(inout) :: self
        type(hip_matrix_t), intent(in) :: A
        real(c_double), intent(in) :: v(:)
        real(c_double), allocatable :: z(:)
        allocate(z(size(v)))
        call hip_check(cdg_prec_apply(A%ctx, self%h, A%h, v, z), 'cdg_prec_apply', A%ctx)
    end function

    subroutine prec_set_schedule(self, schedule)
        class(hip_precon_t), intent(inout) :: self
        integer, intent(in) :: schedule
        call hip_check(cdg_prec_set_schedule(self%ctx, self%h, int(schedule, c_int32_t)), 'cdg_prec_set_schedule', &
                       self%ctx)
    end subroutine

    !> Times apply launch shapes on the current factor (reps calls each) and keeps
    !! the fastest for this preconditioner and later ones of the same block size
    !! and precision; returns its ms per call.  Needs a factor (update first).
    function prec_tune(self, A, reps) result(best_ms)
        class(hip_precon_t), intent(inout) :: self
        type(hip_matrix_t), intent(in) :: A
        integer, intent(in) :: reps
        real(c_double) :: best_ms
        best_ms = 0.d0
        call hip_check(cdg_prec_tune(A%ctx, self%h, A%h, int(reps, c_int32_t), best_ms), 'cdg_prec_tune', A%ctx)
    end function

    subroutine prec_tear_down(self)
        class(hip_precon_t), intent(inout) :: self
        integer(c_int) :: st
        if (c_associated(self%h)) st = cdg_prec_destroy(self%ctx, self%h)
        self%h = c_null_ptr
        self%initialized = .false.
    end subroutine

    subroutine fgmres_init(self)
        class(hip_fgmres_t), intent(inout) :: self
        call cdg_fgmres_default_params(self%params)
    end subroutine

    !> solve(A,x,b,M[,solver_controller]): x overwritten (zero initial guess).
    !! Without a controller M is updated first, as in the reference; with one,
    !! when solver_controller%update_preconditioner says so
    !! (type_fgmres.f90:118-122).
    subroutine fgmres_solve(self, A, x, b, M, solver_controller)
        class(hip_fgmres_t), intent(inout) :: self
        type(hip_matrix_t), intent(in) :: A
        real(c_double), intent(inout) :: x(:)
        real(c_double), intent(in) :: b(:)
        type(hip_precon_t), intent(inout) :: M
        type(hip_solver_controller_t), intent(inout), optional :: solver_controller
        if (.not. M%initialized) call M%init(A)
        self%params%update_preconditioner = 1
        if (present(solver_controller)) then
            self%params%update_preconditioner = 0
            if (solver_controller%update_preconditioner(A, M)) self%params%update_preconditioner = 1
        end if
        call hip_check(cdg_fgmres_solve(A%ctx, A%h, M%h, b, x, self%params, self%stats), 'cdg_fgmres_solve', A%ctx)
        self%niter = self%stats%niter
        self%residual_norm = self%stats%residual_norm
        if (self%stats%factored /= 0) M%stamp = A%stamp()
    end subroutine

    !> solve(A,x,b,M[,solver_controller]) with x and b on the device.
    subroutine fgmres_solve_dev(self, A, x, b, M, solver_controller)
        class(hip_fgmres_t), intent(inout) :: self
        type(hip_matrix_t), intent(in) :: A
        type(hip_vector_t), intent(inout) :: x
        type(hip_vector_t), intent(in) :: b
        type(hip_precon_t), intent(inout) :: M
        type(hip_solver_controller_t), intent(inout), optional :: solver_controller
        real(c_double), pointer, contiguous :: xp(:), bp(:)
        ! the device addresses, passed through as they are (never dereferenced here)
        call c_f_pointer(x%dev, xp, [x%len])
        call c_f_pointer(b%dev, bp, [b%len])
        if (.not. M%initialized) call M%init(A)
        self%params%update_preconditioner = 1
        if (present(solver_controller)) then
            self%params%update_preconditioner = 0
            if (solver_controller%update_preconditioner(A, M)) self%params%update_preconditioner = 1
        end if
        call hip_check(cdg_fgmres_solve(A%ctx, A%h, M%h, bp, xp, self%params, self%stats), 'cdg_fgmres_solve', A%ctx)
        self%niter = self%stats%niter
        self%residual_norm = self%stats%residual_norm
        if (self%stats%factored /= 0) M%stamp = A%stamp()
    end subroutine

    subroutine vec_init(self, ctx, len)
        class(hip_vector_t), intent(inout) :: self
        type(hip_context_t), intent(in) :: ctx
        integer, intent(in) :: len
        self%ctx = ctx%h
        self%len = int(len, c_int64_t)
        call hip_check(cdg_vec_create(self%ctx, self%len, self%dev), 'cdg_vec_create', self%ctx)
    end subroutine

    subroutine vec_set(self, x)
        class(hip_vector_t), intent(inout) :: self
        real(c_double), intent(in), target, contiguous :: x(:)
        call hip_check(cdg_vec_copy(self%ctx, self%len, c_loc(x), self%dev), 'cdg_vec_copy', self%ctx)
    end subroutine

    subroutine vec_get(self, x)
        class(hip_vector_t), intent(in) :: self
        real(c_double), intent(inout), target, contiguous :: x(:)
        call hip_check(cdg_vec_copy(self%ctx, self%len, self%dev, c_loc(x)), 'cdg_vec_copy', self%ctx)
    end subroutine

    subroutine vec_axpby(self, a, x, b)
        class(hip_vector_t), intent(inout) :: self
        real(c_double), intent(in) :: a, b
        type(hip_vector_t), intent(in) :: x
        call hip_check(cdg_vec_axpby(self%ctx, self%len, a, x%dev, b, self%dev), 'cdg_vec_axpby', self%ctx)
    end subroutine

    function vec_norm(self) result(r)
        class(hip_vector_t), intent(in) :: self
        real(c_double) :: r
        call hip_check(cdg_vec_norm(self%ctx, self%len, self%dev, r), 'cdg_vec_norm', self%ctx)
    end function

    function vec_dot(self, x) result(r)
        class(hip_vector_t), intent(in) :: self
        type(hip_vector_t), intent(in) :: x
        real(c_double) :: r
        call hip_check(cdg_vec_dot(self%ctx, self%len, self%dev, x%dev, r), 'cdg_vec_dot', self%ctx)
    end function

    subroutine vec_release(self)
        class(hip_vector_t), intent(inout) :: self
        integer(c_int) :: st
        if (c_associated(self%dev)) st = cdg_vec_destroy(self%ctx, self%dev)
        self%dev = c_null_ptr
    end subroutine

    !> R(q) of the DG Euler discretization on the periodic N^3 box, degree
    !! order, on the device (cdg_euler_residual_gpu; the reference assembles
    !! its residual in mod_spatial.f90:40-197).
    subroutine hip_euler_residual(ctx, N, order, q, R)
        type(hip_context_t), intent(in) :: ctx
        integer, intent(in) :: N, order
        type(hip_vector_t), intent(in) :: q
        type(hip_vector_t), intent(inout) :: R
        call hip_check(cdg_euler_residual_gpu(ctx%h, int(N, c_int32_t), int(N, c_int32_t), int(N, c_int32_t), &
                                              int(order, c_int32_t), q%dev, R%dev), 'cdg_euler_residual_gpu', ctx%h)
    end subroutine

    !> Default: always update the lhs (type_solver_controller.f90:58-74).
    function ctl_update_lhs(self, A, niter, residual_ratio) result(update)
        class(hip_solver_controller_t), intent(inout) :: self
        type(hip_matrix_t), intent(in) :: A
        integer, intent(in) :: niter
        real(c_double), intent(in) :: residual_ratio
        logical :: update
        update = .true.
        self%lhs_updated = .true.
    end function

    !> type_solver_controller.f90:89-124.  The reference reduces the decision
    !! with a logical OR over the ranks of its communicator; one GPU is one
    !! rank, whose decision that reduction returns unchanged.
    function ctl_update_preconditioner(self, lhs, preconditioner) result(update)
        class(hip_solver_controller_t), intent(inout) :: self
        type(hip_matrix_t), intent(in) :: lhs
        type(hip_precon_t), intent(in) :: preconditioner
        logical :: update
        integer(c_int64_t) :: stamp
        stamp = lhs%stamp()
        update = (self%lhs_stamp /= stamp) .or. self%force_update_preconditioner .or. &
                 (stamp /= preconditioner%stamp)
        self%lhs_stamp = stamp
        self%force_update_preconditioner = .false.
        self%preconditioner_updated = update
    end function

    ! ---- bridging from the reference containers' layout ---------------------

    !> row_ptr/col/kind (0-based, C) of rows(:) in ChiDG row order: per row the
    !! lblks entries in their stored order (face order, then the diagonal,
    !! type_domain_matrix.f90:134,253-305), then the coupling entries.
    subroutine hip_pattern_from_rows(rows, row_ptr, col, kind)
        type(hip_block_row_t), intent(in) :: rows(:)
        integer(c_int32_t), allocatable, intent(out) :: row_ptr(:), col(:), kind(:)
        integer :: i, k, nb, nl, nc
        allocate(row_ptr(size(rows) + 1))
        row_ptr(1) = 0
        do i = 1, size(rows)
            row_ptr(i + 1) = row_ptr(i) + nentries(rows(i), nl, nc)
        end do
        nb = row_ptr(size(rows) + 1)
        allocate(col(max(nb, 1)), kind(max(nb, 1)))
        do i = 1, size(rows)
            nb = row_ptr(i) + nentries(rows(i), nl, nc)
            do k = 1, nl
                col(row_ptr(i) + k) = rows(i)%lparent(k) - 1
                kind(row_ptr(i) + k) = HIP_KIND_LOCAL
            end do
            do k = 1, nc
                col(row_ptr(i) + nl + k) = rows(i)%cparent(k) - 1
                kind(row_ptr(i) + nl + k) = HIP_KIND_COUPLING
            end do
        end do
    end subroutine

    !> All blocks of rows(:) concatenated in pattern order, each n x n
    !! column-major (densematrix_t%mat's own storage order), ready for store.
    subroutine hip_pack_blocks(rows, n, blocks)
        type(hip_block_row_t), intent(in) :: rows(:)
        integer, intent(in) :: n
        real(c_double), allocatable, intent(out) :: blocks(:)
        integer :: i, k, nl, nc, nb, pos
        nb = 0
        do i = 1, size(rows)
            nb = nb + nentries(rows(i), nl, nc)
        end do
        allocate(blocks(max(nb, 1) * n * n))
        pos = 0
        do i = 1, size(rows)
            nb = nentries(rows(i), nl, nc)
            do k = 1, nl
                blocks(pos + 1:pos + n * n) = reshape(rows(i)%lmat(:, :, k), [n * n])
                pos = pos + n * n
            end do
            do k = 1, nc
                blocks(pos + 1:pos + n * n) = reshape(rows(i)%cmat(:, :, k), [n * n])
                pos = pos + n * n
            end do
        end do
    end subroutine

    !> x = the element vectors of a domain_vector_t%vecs(:) one after another
    !! (element-major; within an element the reference's (ifield-1)*nterms+iterm
    !! order, type_densevector.f90:214).
    subroutine hip_get_packed(vecs, x)
        type(hip_densevector_t), intent(in) :: vecs(:)
        real(c_double), allocatable, intent(out) :: x(:)
        integer :: i, pos, total
        total = 0
        do i = 1, size(vecs)
            total = total + size(vecs(i)%vec)
        end do
        allocate(x(total))
        pos = 0
        do i = 1, size(vecs)
            x(pos + 1:pos + size(vecs(i)%vec)) = vecs(i)%vec
            pos = pos + size(vecs(i)%vec)
        end do
    end subroutine

    !> The inverse of hip_get_packed: x scattered back into the (allocated)
    !! element vectors.
    subroutine hip_set_packed(x, vecs)
        real(c_double), intent(in) :: x(:)
        type(hip_densevector_t), intent(inout) :: vecs(:)
        integer :: i, pos
        pos = 0
        do i = 1, size(vecs)
            vecs(i)%vec = x(pos + 1:pos + size(vecs(i)%vec))
            pos = pos + size(vecs(i)%vec)
        end do
    end subroutine

    integer function nentries(row, nl, nc)
        type(hip_block_row_t), intent(in) :: row
        integer, intent(out) :: nl, nc
        nl = 0
        nc = 0
        if (allocated(row%lparent)) nl = size(row%lparent)
        if (allocated(row%cparent)) nc = size(row%cparent)
        nentries = nl + nc
    end function

end module mod_chidg_hip
