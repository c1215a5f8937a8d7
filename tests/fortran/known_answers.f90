!> Reference known answers through the Fortran binding (mod_chidg_hip):
!!   test/unit_testing/preconditioners/test_ILU0_mod.pf:227-332  (ILU0 factor + apply)
!!   test/unit_testing/linear_solvers/test_fgmres.pf             (FGMRES 3x3, identity)
!!   test/unit_testing/linear_solvers/test_fgmres_ref.pf         (FGMRES 2x2 skew)
!! The 3- and 2-element D1 meshes give 1x1 blocks in ChiDG's lblks order
!! [XI_MIN, XI_MAX, DIAG].  Prints PASS/FAIL lines and "ALL PASS".
program known_answers
    use iso_c_binding
    use mod_chidg_hip
    implicit none
    type(hip_context_t) :: ctx
    type(hip_matrix_t)  :: A, B
    type(hip_precon_t)  :: M, I3, I2
    type(hip_fgmres_t)  :: ls
    type(hip_matrix_t)  :: C
    type(hip_precon_t)  :: MC
    type(hip_solver_controller_t) :: ctl
    type(hip_block_row_t) :: rows(3)
    type(hip_densevector_t) :: vecs(3)
    integer(c_int32_t), allocatable :: rp(:), cl(:), kd(:)
    real(c_double), allocatable :: z(:), x(:), packed(:), xp(:)
    real(c_double) :: blk(7), t_apply, t_mv
    integer :: nfail = 0, k, factored(5)

    call ctx%init(0)

    ! --- ILU0: A = [[1,-2,0],[2,-5,12],[0,2,-10]], b = [1,2,3] ---------------
    call A%init(ctx, 3, 1, [0, 2, 5, 7], [1, 0, 0, 2, 1, 1, 2], [0, 0, 0, 0, 0, 0, 0])
    call A%store([-2.d0, 1.d0, 2.d0, 12.d0, -5.d0, 2.d0, -10.d0])
    call A%fetch(blk)
    call check('mat_fetch', all(blk == [-2.d0, 1.d0, 2.d0, 12.d0, -5.d0, 2.d0, -10.d0]))
    call check('mem_used', ctx%mem_used() > 0)
    call M%init(A, HIP_PREC_ILU0, HIP_FP64)
    call M%update(A)
    z = M%apply(A, [1.d0, 2.d0, 3.d0])
    call check('ilu0_apply', maxval(abs(z - [43.d0/7.d0, 18.d0/7.d0, 3.d0/14.d0])) < 1.d-12)
    call check('ilu0_apply_reference_tol', abs(z(1) - 6.142857143d0) < 1.d-5 .and. &
               abs(z(2) - 2.57142857d0) < 1.d-5 .and. abs(z(3) - 0.2142857143d0) < 1.d-5)
    ! the same answer with Dinv folded into the upper blocks, then under tuned launch shapes
    call M%set_schedule(HIP_SCHEDULE_FOLDED)
    z = M%apply(A, [1.d0, 2.d0, 3.d0])
    call check('ilu0_apply_folded', maxval(abs(z - [43.d0/7.d0, 18.d0/7.d0, 3.d0/14.d0])) < 1.d-12)
    call M%set_schedule(HIP_SCHEDULE_LEVELS)
    t_apply = M%tune(A, 2)
    t_mv = A%tune(2)
    z = M%apply(A, [1.d0, 2.d0, 3.d0])
    call check('ilu0_apply_tuned', t_apply > 0.d0 .and. t_mv > 0.d0 .and. &
               maxval(abs(z - [43.d0/7.d0, 18.d0/7.d0, 3.d0/14.d0])) < 1.d-12)
    call M%set_schedule(HIP_SCHEDULE_STREAM)
    z = M%apply(A, [1.d0, 2.d0, 3.d0])
    call check('ilu0_apply_stream', maxval(abs(z - [43.d0/7.d0, 18.d0/7.d0, 3.d0/14.d0])) < 1.d-12)
    call M%set_schedule(HIP_SCHEDULE_AUTO)

    ! --- the same system from rows in domain_matrix_t layout (lblks parents in
    !     face order [XI_MIN, XI_MAX, DIAG], 1-based) through the bridging ----
    allocate(rows(1)%lparent(2), rows(1)%lmat(1, 1, 2))
    rows(1)%lparent = [2, 1]
    rows(1)%lmat(1, 1, :) = [-2.d0, 1.d0]
    allocate(rows(2)%lparent(3), rows(2)%lmat(1, 1, 3))
    rows(2)%lparent = [1, 3, 2]
    rows(2)%lmat(1, 1, :) = [2.d0, 12.d0, -5.d0]
    allocate(rows(3)%lparent(2), rows(3)%lmat(1, 1, 2), rows(3)%cparent(1), rows(3)%cmat(1, 1, 1))
    rows(3)%lparent = [2, 3]
    rows(3)%lmat(1, 1, :) = [2.d0, -10.d0]
    rows(3)%cparent = [1]            ! a coupling-only (chi_blks) entry: matvec only
    rows(3)%cmat(1, 1, 1) = 0.5d0
    call hip_pattern_from_rows(rows, rp, cl, kd)
    call check('bridge_pattern', all(rp == [0, 2, 5, 8]) .and. all(cl == [1, 0, 0, 2, 1, 1, 2, 0]) .and. &
               all(kd == [0, 0, 0, 0, 0, 0, 0, 1]))
    call hip_pack_blocks(rows, 1, packed)
    call check('bridge_pack', all(packed == [-2.d0, 1.d0, 2.d0, 12.d0, -5.d0, 2.d0, -10.d0, 0.5d0]))
    call C%init(ctx, 3, 1, rp, cl, kd)
    call C%store(packed)
    do k = 1, 3
        allocate(vecs(k)%vec(1))
        vecs(k)%vec = real(k, c_double)
    end do
    call hip_get_packed(vecs, xp)
    call check('bridge_get_packed', all(xp == [1.d0, 2.d0, 3.d0]))
    z = hip_mv(C, xp)                ! lblks product plus 0.5 * x(1) into row 3
    call check('bridge_mv', maxval(abs(z - [-3.d0, 28.d0, -25.5d0])) < 1.d-13)
    call hip_set_packed(z, vecs)
    call check('bridge_set_packed', vecs(2)%vec(1) == 28.d0 .and. vecs(3)%vec(1) == -25.5d0)

    ! --- solver_controller: refactor only when the lhs changed or when forced -
    call MC%init(C, HIP_PREC_ILU0, HIP_FP64)
    call ls%init()
    ls%params%inner_fgmres = 0
    ls%params%tol = 1.d-13
    allocate(x(3))
    do k = 1, 5
        if (k == 3) call C%store(packed)                 ! a new lhs version (re-stamped)
        if (k == 5) ctl%force_update_preconditioner = .true.
        call ls%solve(C, x, [1.d0, 2.d0, 3.d0], MC, ctl)
        factored(k) = ls%stats%factored
    end do
    call check('controller_refactor_rule', all(factored == [1, 0, 1, 0, 1]))
    call check('controller_solution', maxval(abs(hip_mv(C, x) - [1.d0, 2.d0, 3.d0])) < 1.d-10)

    ! --- FGMRES 3x3: A = [[4,3,0],[3,4,-1],[0,-1,4]], b = [24,30,-24] --------
    call A%store([3.d0, 4.d0, 3.d0, -1.d0, 4.d0, -1.d0, 4.d0])
    call ls%init()                 ! reference defaults (inner FGMRES on)
    ls%params%tol = 1.d-13
    call I3%init(A, HIP_PREC_IDENTITY)
    call ls%solve(A, x, [24.d0, 30.d0, -24.d0], I3)
    call check('fgmres_3x3', maxval(abs(x - [3.d0, 4.d0, -5.d0])) < 1.d-9)

    ! --- FGMRES 2x2 skew: A = [[0,1],[-1,0]], b = [1,1] ------------------------
    call B%init(ctx, 2, 1, [0, 2, 4], [1, 0, 0, 1], [0, 0, 0, 0])
    call B%store([1.d0, 0.d0, -1.d0, 0.d0])
    call I2%init(B, HIP_PREC_IDENTITY)
    deallocate(x)
    allocate(x(2))
    call ls%solve(B, x, [1.d0, 1.d0], I2)
    call check('fgmres_2x2_skew', maxval(abs(x - [-1.d0, 1.d0])) < 1.d-9)

    ! --- matvec through the binding ---------------------------------------------
    z = hip_mv(A, [3.d0, 4.d0, -5.d0])
    call check('mv', maxval(abs(z - [24.d0, 30.d0, -24.d0])) < 1.d-13)

    call M%tear_down()
    call MC%tear_down()
    call C%release()
    call I3%tear_down()
    call I2%tear_down()
    call A%release()
    call B%release()
    call ctx%release()
    if (nfail == 0) then
        print '(a)', 'ALL PASS'
    else
        print '(a, i0)', 'FAILURES: ', nfail
        error stop 1
    end if

contains

    subroutine check(name, ok)
        character(*), intent(in) :: name
        logical, intent(in) :: ok
        if (ok) then
            print '(a, a)', 'PASS ', name
        else
            print '(a, a)', 'FAIL ', name
            nfail = nfail + 1
        end if
    end subroutine

end program known_answers
