!> Pseudo-transient Newton steps on the 4^3 P1 DG Euler box with q, dq and
!! rhs resident on the device (hip_vector_t), the way a ChiDG build with
!! backend 'hip' keeps them across Newton steps (type_newton.f90:64-266):
!! rhs = -R(q) on the device (hip_euler_residual), the Jacobian + mass/dtau
!! assembled on the host from q and stored (chidg_matrix%assemble),
!! solve(A, dq, rhs, M) with device vectors, q = q + dq on the device.  The
!! same steps with host arrays and the host residual must give the same
!! residual norms (1e-10) and q (1e-12).  Prints PASS/FAIL lines and "ALL PASS".
program newton_device
    use iso_c_binding
    use mod_chidg_hip
    implicit none
    interface
        integer(c_int) function cdg_euler_state(nx, ny, nz, order, mach, seed, perturb, q) &
                bind(c, name='cdg_euler_state')
            import :: c_int, c_int32_t, c_int64_t, c_double
            integer(c_int32_t), value :: nx, ny, nz, order
            real(c_double), value :: mach, perturb
            integer(c_int64_t), value :: seed
            real(c_double) :: q(*)
        end function
        integer(c_int) function cdg_euler_jacobian(nx, ny, nz, order, q, cfl, blocks, R, nthreads) &
                bind(c, name='cdg_euler_jacobian')
            import :: c_int, c_int32_t, c_double
            integer(c_int32_t), value :: nx, ny, nz, order, nthreads
            real(c_double), intent(in) :: q(*)
            real(c_double), value :: cfl
            real(c_double) :: blocks(*), R(*)
        end function
    end interface
    integer, parameter :: NE = 4, ORD = 1, NB = 40, NROWS = NE**3, LEN = NROWS * NB, NSTEPS = 3
    type(hip_context_t) :: ctx
    type(hip_matrix_t) :: A
    type(hip_precon_t) :: M
    type(hip_fgmres_t) :: ls
    type(hip_vector_t) :: q, dq, rhs
    integer(c_int32_t), allocatable :: rp(:), cl(:), kd(:)
    real(c_double), allocatable :: q0(:), qh(:), xh(:), Rh(:), blocks(:), qt(:)
    real(c_double) :: norms_h(NSTEPS), norms_d(NSTEPS)
    integer :: step, nfail = 0

    call hex_box(NE, rp, cl, kd)
    allocate(q0(LEN), qh(LEN), xh(LEN), Rh(LEN), qt(LEN), blocks(int(rp(NROWS + 1)) * NB * NB))
    call ensure(cdg_euler_state(NE, NE, NE, ORD, 0.3d0, 1_c_int64_t, 0.05d0, q0), 'cdg_euler_state')

    call ctx%init(0)
    call A%init(ctx, NROWS, NB, rp, cl, kd)
    call M%init(A, HIP_PREC_ILU0, HIP_FP64)
    call ls%init()
    ls%params%nkrylov = 30
    ls%params%tol = 0.d0
    ls%params%rtol = 1.d-10
    ls%params%inner_fgmres = 0

    ! host vectors: R on the host, dq and rhs as Fortran arrays
    qh = q0
    do step = 1, NSTEPS
        call ensure(cdg_euler_jacobian(NE, NE, NE, ORD, qh, 1.d0, blocks, Rh, 0), 'cdg_euler_jacobian')
        call A%store(blocks)
        norms_h(step) = norm2(Rh)
        call ls%solve(A, xh, -Rh, M)
        qh = qh + xh
    end do

    ! device vectors: q, dq, rhs stay on the device; only q goes to the host,
    ! where the Jacobian is assembled
    call q%init(ctx, LEN)
    call dq%init(ctx, LEN)
    call rhs%init(ctx, LEN)
    call q%set(q0)
    do step = 1, NSTEPS
        call hip_euler_residual(ctx, NE, ORD, q, rhs)
        norms_d(step) = rhs%norm()
        call rhs%axpby(0.d0, rhs, -1.d0)                  ! rhs = -R
        call q%get(qt)
        call ensure(cdg_euler_jacobian(NE, NE, NE, ORD, qt, 1.d0, blocks, Rh, 0), 'cdg_euler_jacobian')
        call A%store(blocks)
        call ls%solve_dev(A, dq, rhs, M)
        call q%axpby(1.d0, dq, 1.d0)                      ! q = q + dq
    end do
    call q%get(qt)

    call check('newton_device_residual_norms', maxval(abs(norms_d - norms_h) / norms_h) < 1.d-10)
    call check('newton_device_residual_falls', norms_d(3) < norms_d(2) .and. norms_d(2) < norms_d(1))
    call check('newton_device_state', norm2(qt - qh) / norm2(qh) < 1.d-12)
    call check('device_vector_dot', abs(q%dot(q) - dot_product(qh, qh)) < 1.d-10 * dot_product(qh, qh))
    print '(a, 3es12.4)', 'residual norms (device):', norms_d

    call q%release()
    call dq%release()
    call rhs%release()
    call M%tear_down()
    call A%release()
    call ctx%release()
    if (nfail == 0) then
        print '(a)', 'ALL PASS'
    else
        print '(a, i0, a)', 'FAILED ', nfail, ' check(s)'
        stop 1
    end if

contains

    subroutine check(name, ok)
        character(*), intent(in) :: name
        logical, intent(in) :: ok
        if (ok) then
            print '(2a)', 'PASS ', name
        else
            print '(2a)', 'FAIL ', name
            nfail = nfail + 1
        end if
    end subroutine

    subroutine ensure(st, what)
        integer(c_int), intent(in) :: st
        character(*), intent(in) :: what
        if (st /= 0) then
            print '(3a, i0)', 'FAIL ', what, ' status ', st
            stop 1
        end if
    end subroutine

    !> The periodic nx^3 hex box in ChiDG row order: per element the interior
    !! neighbours in face order XI_MIN, XI_MAX, ETA_MIN, ETA_MAX, ZETA_MIN,
    !! ZETA_MAX, then the diagonal (lblks), then the periodic donors as
    !! coupling-only entries; 0-based row_ptr and parents.
    subroutine hex_box(nx, row_ptr, col, kind)
        integer, intent(in) :: nx
        integer(c_int32_t), allocatable, intent(out) :: row_ptr(:), col(:), kind(:)
        integer :: e, d, side, k, qq, wrap, donor, nchi, c, nrows
        integer :: idx(3), stride(3), chi(6)
        nrows = nx**3
        stride = [1, nx, nx * nx]
        allocate(row_ptr(nrows + 1), col(7 * nrows), kind(7 * nrows))
        row_ptr(1) = 0
        k = 0
        do e = 0, nrows - 1
            idx = [mod(e, nx), mod(e / nx, nx), e / (nx * nx)]
            nchi = 0
            do d = 1, 3
                do side = -1, 1, 2
                    qq = idx(d) + side
                    if (qq >= 0 .and. qq < nx) then
                        k = k + 1
                        col(k) = e + side * stride(d)
                        kind(k) = HIP_KIND_LOCAL
                    else
                        wrap = mod(qq + nx, nx)
                        donor = e + (wrap - idx(d)) * stride(d)
                        if (.not. any(chi(1:nchi) == donor)) then
                            nchi = nchi + 1
                            chi(nchi) = donor
                        end if
                    end if
                end do
            end do
            k = k + 1
            col(k) = e
            kind(k) = HIP_KIND_LOCAL
            do c = 1, nchi
                k = k + 1
                col(k) = chi(c)
                kind(k) = HIP_KIND_COUPLING
            end do
            row_ptr(e + 2) = k
        end do
    end subroutine
end program newton_device
