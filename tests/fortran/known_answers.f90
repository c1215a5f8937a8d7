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
    real(c_double), allocatable :: z(:), x(:)
    real(c_double) :: blk(7), t_apply, t_mv
    integer :: nfail = 0

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
    call M%set_schedule(HIP_SCHEDULE_AUTO)

    ! --- FGMRES 3x3: A = [[4,3,0],[3,4,-1],[0,-1,4]], b = [24,30,-24] --------
    call A%store([3.d0, 4.d0, 3.d0, -1.d0, 4.d0, -1.d0, 4.d0])
    call ls%init()                 ! reference defaults (inner FGMRES on)
    ls%params%tol = 1.d-13
    call I3%init(A, HIP_PREC_IDENTITY)
    allocate(x(3))
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
