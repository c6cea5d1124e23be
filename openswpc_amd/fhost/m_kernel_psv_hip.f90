!! Drop-in replacements for the P-SV reference's parameterless hot-path
!! routines, backed by the MI355X library through module swpc_hip.
!!
!! The reference's main loop (src/swpc_psv/main.f90:95-114) calls
!! snap__write(it), wav__store(it), kernel__update_stress,
!! absorb__update_stress, source__stressglut(it), kernel__update_vel,
!! source__bodyforce(it) and absorb__update_vel on module-global arrays.
!! One context, created by kernel_psv_hip__setup from the same arrays
!! (k,i layout over the memory box [-2:n+3]), holds their device copies.
!! The library updates interior and PML cells in one call, so the absorb__*
!! routines only keep the call sites valid.
module m_kernel_psv_hip

    use iso_c_binding
    use iso_fortran_env, only: error_unit
    use swpc_hip
    implicit none
    private

    type(c_ptr), save :: ctx = c_null_ptr

    public :: kernel_psv_hip__setup, kernel_psv_hip__finish, kernel_psv_hip__check, kernel_psv_hip__ctx
    public :: kernel__update_stress, kernel__update_vel, absorb__update_stress, absorb__update_vel
    public :: source__stressglut, source__bodyforce, wav__store, snap__write, kernel__vmax, wav__get

contains

    !! Stop as the reference's assert macro does (m_pdebug.f90:223-240),
    !! with the library's message.
    subroutine kernel_psv_hip__check(ierr, what)
        integer(c_int), intent(in) :: ierr
        character(*), intent(in) :: what
        if (ierr /= SWPC_OK) then
            write (error_unit, '(A,I0,2A)') 'swpc_hip: '//what//' failed (', ierr, '): ', swpc_errmsg()
            error stop 1
        end if
    end subroutine kernel_psv_hip__check

    function kernel_psv_hip__ctx() result(c)
        type(c_ptr) :: c
        c = ctx
    end function kernel_psv_hip__ctx

    !! kernel__setup + absorb__setup in one: rho, lam, mu, taup, taus after
    !! medium__setup (src/swpc_psv/m_medium.f90), the surface maps over i and
    !! the damping profiles g(4, n) of absorb_p__setup.
    subroutine kernel_psv_hip__setup(desc, rho, lam, mu, taup, taus, kfs_top, kfs_bot, kob_top, kob_bot, kob, &
                                     gxc, gxe, gzc, gze)
        type(swpc2d_desc), intent(in) :: desc
        real(c_float), intent(in) :: rho(*), lam(*), mu(*), taup(*), taus(*)
        integer(c_int), intent(in) :: kfs_top(*), kfs_bot(*), kob_top(*), kob_bot(*), kob(*)
        real(c_float), intent(in) :: gxc(*), gxe(*), gzc(*), gze(*)
        call kernel_psv_hip__check(swpc2d_create(desc, ctx), 'swpc2d_create')
        call kernel_psv_hip__check(swpc2d_set_medium(ctx, rho, lam, mu, taup, taus), 'set_medium')
        call kernel_psv_hip__check(swpc2d_set_surface(ctx, kfs_top, kfs_bot, kob_top, kob_bot, kob), 'set_surface')
        call kernel_psv_hip__check(swpc2d_set_pml(ctx, gxc, gxe, gzc, gze), 'set_pml')
    end subroutine kernel_psv_hip__setup

    subroutine kernel_psv_hip__finish()
        integer(c_int) :: ierr
        if (c_associated(ctx)) ierr = swpc2d_destroy(ctx)
        ctx = c_null_ptr
    end subroutine kernel_psv_hip__finish

    subroutine kernel__update_stress()
        call kernel_psv_hip__check(swpc2d_update_stress(ctx), 'update_stress')
    end subroutine

    subroutine kernel__update_vel()
        call kernel_psv_hip__check(swpc2d_update_vel(ctx), 'update_vel')
    end subroutine

    subroutine absorb__update_stress()
        !! the PML cells are updated inside kernel__update_stress
    end subroutine

    subroutine absorb__update_vel()
        !! the PML cells are updated inside kernel__update_vel
    end subroutine

    subroutine source__stressglut(it)
        integer, intent(in) :: it
        call kernel_psv_hip__check(swpc2d_stressglut(ctx, int(it, c_int)), 'stressglut')
    end subroutine

    subroutine source__bodyforce(it)
        integer, intent(in) :: it
        call kernel_psv_hip__check(swpc2d_bodyforce(ctx, int(it, c_int)), 'bodyforce')
    end subroutine

    subroutine wav__store(it)
        integer, intent(in) :: it
        call kernel_psv_hip__check(swpc2d_store_receivers(ctx, int(it, c_int)), 'store_receivers')
    end subroutine

    subroutine snap__write(it)
        integer, intent(in) :: it
        call kernel_psv_hip__check(swpc2d_snapshots(ctx, int(it, c_int)), 'snapshots')
    end subroutine

    !! kernel__vmax (src/swpc_psv/m_kernel.f90:308-331)
    subroutine kernel__vmax(vx, vz)
        real, intent(out) :: vx, vz
        real(c_float) :: v(2)
        call kernel_psv_hip__check(swpc2d_vmax(ctx, v), 'vmax')
        vx = v(1); vz = v(2)
    end subroutine

    !! receiver traces gathered to the host: vel(2, nst, ntw), disp(2, nst, ntw)
    subroutine wav__get(vel, disp)
        real(c_float), intent(out) :: vel(*)
        real(c_double), intent(out) :: disp(*)
        call kernel_psv_hip__check(swpc2d_get_receivers(ctx, vel, disp), 'get_receivers')
    end subroutine

end module m_kernel_psv_hip
