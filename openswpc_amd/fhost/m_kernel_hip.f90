!! Drop-in replacements for the reference's parameterless hot-path
!! routines, backed by the MI355X library through module swpc_hip.
!!
!! The reference's main loop (src/swpc_3d/main.f90:125-144) calls
!! kernel__update_stress, absorb__update_stress, source__stressglut(it),
!! kernel__update_vel, absorb__update_vel, source__bodyforce(it) and
!! wav__store(it) on module-global arrays.  Here one context, created by
!! kernel_hip__setup from the same arrays, holds their device copies.  The
!! library updates interior and PML cells in one call, so the absorb__*
!! routines only keep the call sites valid.
module m_kernel_hip

    use iso_c_binding
    use iso_fortran_env, only: error_unit
    use swpc_hip
    implicit none
    private

    type(c_ptr), save :: ctx = c_null_ptr

    public :: kernel_hip__setup, kernel_hip__finish, kernel_hip__check
    public :: kernel__update_stress, kernel__update_vel, absorb__update_stress, absorb__update_vel
    public :: source__stressglut, source__bodyforce, wav__store, snap__write, kernel__vmax
    public :: wav__get, kernel_hip__ctx

contains

    !! Stop as the reference's assert macro does (m_pdebug.f90:223-240),
    !! with the library's message.
    subroutine kernel_hip__check(ierr, what)
        integer(c_int), intent(in) :: ierr
        character(*), intent(in) :: what
        if (ierr /= SWPC_OK) then
            write (error_unit, '(A,I0,2A)') 'swpc_hip: '//what//' failed (', ierr, '): ', swpc_errmsg()
            error stop 1
        end if
    end subroutine kernel_hip__check

    function kernel_hip__ctx() result(c)
        type(c_ptr) :: c
        c = ctx
    end function kernel_hip__ctx

    !! kernel__setup + absorb__setup + source__setup + wav__setup in one:
    !! arrays in the reference layout over the memory box.
    subroutine kernel_hip__setup(desc, bx, by, bz, lam, mu, taup, taus, muyz, muxz, muxy, &
                                 kfs_top, kfs_bot, kob_top, kob_bot, kob, gxc, gxe, gyc, gye, gzc, gze)
        type(swpc3d_desc), intent(in) :: desc
        real(c_float), intent(in) :: bx(*), by(*), bz(*), lam(*), mu(*), taup(*), taus(*), muyz(*), muxz(*), muxy(*)
        integer(c_int), intent(in) :: kfs_top(*), kfs_bot(*), kob_top(*), kob_bot(*), kob(*)
        real(c_float), intent(in) :: gxc(*), gxe(*), gyc(*), gye(*), gzc(*), gze(*)
        call kernel_hip__check(swpc3d_create(desc, ctx), 'swpc3d_create')
        call kernel_hip__check(swpc3d_set_medium(ctx, bx, by, bz, lam, mu, taup, taus, muyz, muxz, muxy), &
                               'set_medium')
        call kernel_hip__check(swpc3d_set_surface(ctx, kfs_top, kfs_bot, kob_top, kob_bot, kob), 'set_surface')
        call kernel_hip__check(swpc3d_set_pml(ctx, gxc, gxe, gyc, gye, gzc, gze), 'set_pml')
    end subroutine kernel_hip__setup

    subroutine kernel_hip__finish()
        integer(c_int) :: ierr
        if (c_associated(ctx)) ierr = swpc3d_destroy(ctx)
        ctx = c_null_ptr
    end subroutine kernel_hip__finish

    subroutine kernel__update_stress()
        call kernel_hip__check(swpc3d_update_stress(ctx), 'update_stress')
    end subroutine

    subroutine kernel__update_vel()
        call kernel_hip__check(swpc3d_update_vel(ctx), 'update_vel')
    end subroutine

    subroutine absorb__update_stress()
        !! the PML boxes are updated inside kernel__update_stress
    end subroutine

    subroutine absorb__update_vel()
        !! the PML boxes are updated inside kernel__update_vel
    end subroutine

    subroutine source__stressglut(it)
        integer, intent(in) :: it
        call kernel_hip__check(swpc3d_stressglut(ctx, int(it, c_int)), 'stressglut')
    end subroutine

    subroutine source__bodyforce(it)
        integer, intent(in) :: it
        call kernel_hip__check(swpc3d_bodyforce(ctx, int(it, c_int)), 'bodyforce')
    end subroutine

    subroutine wav__store(it)
        integer, intent(in) :: it
        call kernel_hip__check(swpc3d_store_receivers(ctx, int(it, c_int)), 'store_receivers')
    end subroutine

    subroutine snap__write(it)
        integer, intent(in) :: it
        call kernel_hip__check(swpc3d_snapshots(ctx, int(it, c_int)), 'snapshots')
    end subroutine

    !! kernel__vmax (m_kernel.f90:328-352)
    subroutine kernel__vmax(vx, vy, vz)
        real, intent(out) :: vx, vy, vz
        real(c_float) :: v(3)
        call kernel_hip__check(swpc3d_vmax(ctx, v), 'vmax')
        vx = v(1); vy = v(2); vz = v(3)
    end subroutine

    !! receiver traces gathered to the host: vel(3, nst, ntw), disp(3, nst, ntw)
    subroutine wav__get(vel, disp)
        real(c_float), intent(out) :: vel(*)
        real(c_double), intent(out) :: disp(*)
        call kernel_hip__check(swpc3d_get_receivers(ctx, vel, disp), 'get_receivers')
    end subroutine

end module m_kernel_hip
