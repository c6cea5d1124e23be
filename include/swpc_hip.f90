!! swpc_hip.f90 -- Fortran bind(C) interfaces to include/swpc_hip.h.
!!
!! A Fortran host (the reference's main.f90 and its modules) reaches the
!! MI355X time-stepping core through this module.  Every routine is a
!! function returning 0 on success and a negative code on error; the text
!! of the last error is swpc_errmsg().  Arrays are passed in the
!! reference's own layout, (k,i,j) over the memory box [-2:n+3]
!! (src/swpc_3d/m_global.f90:297-302); an array section such as
!! lam(-2:nz+3, -2:nx+3, -2:ny+3) of a padded allocation is passed as a
!! contiguous copy by the assumed-size dummies below.
module swpc_hip

    use iso_c_binding
    implicit none
    private

    integer(c_int), parameter, public :: SWPC_FP32 = 0, SWPC_FP64 = 1, SWPC_FP32_S16 = 2, SWPC_FP32_R16 = 3
    integer(c_int), parameter, public :: SWPC_OK = 0, SWPC_EINVAL = -1, SWPC_EHIP = -2, SWPC_ESTATE = -3, &
                                         SWPC_ENOMEM = -4
    integer(c_int), parameter, public :: SWPC_SNAP_PS = 0, SWPC_SNAP_V = 1, SWPC_SNAP_U = 2

    type, bind(C), public :: swpc3d_desc
        integer(c_int) :: nx, ny, nz, na, nm, precision
        real(c_double) :: dx, dy, dz, dt
        real(c_float)  :: ts(8)
    end type swpc3d_desc

    type, bind(C), public :: swpc2d_desc
        integer(c_int) :: nx, nz, na, nm, precision
        real(c_double) :: dx, dz, dt
        real(c_float)  :: ts(8)
    end type swpc2d_desc

    public :: swpc_errmsg
    public :: swpc_version, swpc_device_count, swpc_rmedia_generate
    public :: swpc3d_create, swpc3d_destroy, swpc3d_set_medium, swpc3d_set_density, swpc3d_set_surface, swpc3d_set_pml
    public :: swpc3d_setup_medium, swpc3d_get_medium
    public :: swpc3d_set_sources, swpc3d_set_bodyforce, swpc3d_set_receivers
    public :: swpc3d_update_stress, swpc3d_update_vel, swpc3d_stressglut, swpc3d_bodyforce
    public :: swpc3d_store_receivers, swpc3d_add_snapshot, swpc3d_snapshots, swpc3d_get_snapshot
    public :: swpc3d_snapshot_peaks, swpc3d_finish_snapshots, swpc3d_get_peaks
    public :: swpc3d_step, swpc3d_run, swpc3d_get_receivers, swpc3d_vmax, swpc3d_sync
    public :: swpc2d_create, swpc2d_destroy, swpc2d_set_medium, swpc2d_set_surface, swpc2d_set_pml
    public :: swpc2d_set_sources, swpc2d_set_bodyforce, swpc2d_set_receivers, swpc2d_update_stress
    public :: swpc2d_update_vel, swpc2d_stressglut, swpc2d_bodyforce, swpc2d_store_receivers
    public :: swpc2d_add_snapshot, swpc2d_snapshots, swpc2d_get_snapshot
    public :: swpc2d_step, swpc2d_set_graph, swpc2d_set_fusion, swpc2d_run, swpc2d_get_receivers, swpc2d_vmax, swpc2d_sync
    public :: swpcsh_create, swpcsh_destroy, swpcsh_set_medium, swpcsh_set_surface, swpcsh_set_pml
    public :: swpcsh_set_sources, swpcsh_set_bodyforce, swpcsh_set_receivers, swpcsh_update_stress
    public :: swpcsh_update_vel, swpcsh_stressglut, swpcsh_bodyforce, swpcsh_store_receivers
    public :: swpcsh_add_snapshot, swpcsh_snapshots, swpcsh_get_snapshot
    public :: swpcsh_step, swpcsh_set_graph, swpcsh_run, swpcsh_get_receivers, swpcsh_vmax, swpcsh_sync
    public :: swpc_lab_nparam, swpc3d_set_lab, swpc3d_set_lab_sponge
    public :: swpc_batch_create, swpc_batch_destroy, swpc_batch_add3d, swpc_batch_add2d, swpc_batch_run
    public :: swpc_batch_sync, swpc_batch_set_grouping

    interface

        function swpc_last_error() result(p) bind(C, name='swpc_last_error')
            import :: c_ptr
            type(c_ptr) :: p
        end function

        integer(c_int) function swpc_version() bind(C, name='swpc_version')
            import :: c_int
        end function

        integer(c_int) function swpc_device_count() bind(C, name='swpc_device_count')
            import :: c_int
        end function

        ! shape, d, a in C order (last axis fastest): pass them reversed
        ! from Fortran's (nx, ny, nz) so that x stays the fastest axis
        integer(c_int) function swpc_rmedia_generate(nd, shape, d, a, epsil, kappa, ptype, noise, out) &
                bind(C, name='swpc_rmedia_generate')
            import :: c_int, c_double, c_float
            integer(c_int), value :: nd, ptype
            integer(c_int), intent(in) :: shape(*)
            real(c_double), intent(in) :: d(*), a(*)
            real(c_double), value :: epsil, kappa
            real(c_float), intent(in) :: noise(*)
            real(c_float) :: out(*)
        end function

        ! ---------------------------------------------------------------- 3-D

        integer(c_int) function swpc3d_create(desc, ctx) bind(C, name='swpc3d_create')
            import :: c_int, c_ptr, swpc3d_desc
            type(swpc3d_desc), intent(in) :: desc
            type(c_ptr), intent(out) :: ctx
        end function

        integer(c_int) function swpc3d_destroy(ctx) bind(C, name='swpc3d_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        !! medium__setup arrays (m_medium.f90:36-467)
        integer(c_int) function swpc3d_set_medium(ctx, bx, by, bz, lam, mu, taup, taus, muyz, muxz, muxy) &
            bind(C, name='swpc3d_set_medium')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            real(c_float), intent(in) :: bx(*), by(*), bz(*), lam(*), mu(*), taup(*), taus(*)
            real(c_float), intent(in) :: muyz(*), muxz(*), muxy(*)
        end function

        !! optional density over the memory box (interior buoyancy formed in-kernel)
        integer(c_int) function swpc3d_set_density(ctx, rho) bind(C, name='swpc3d_set_density')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            real(c_float), intent(in) :: rho(*)
        end function

        !! medium__setup on the GPU from the velocity model's rho, lam, mu, qp, qs
        integer(c_int) function swpc3d_setup_medium(ctx, rho, lam, mu, qp, qs, zeta3, cc_re, cc_im, kfs, kob, vminmax) &
            bind(C, name='swpc3d_setup_medium')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            real(c_float), intent(in) :: rho(*), lam(*), mu(*), qp(*), qs(*)
            real(c_float), value :: zeta3, cc_re, cc_im
            integer(c_int), intent(out) :: kfs(*), kob(*)
            real(c_float), intent(out) :: vminmax(2)
        end function

        !! one medium array (0..9 set_medium order, 10 rho) back in the reference's order
        integer(c_int) function swpc3d_get_medium(ctx, which, out) bind(C, name='swpc3d_get_medium')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            integer(c_int), value :: which
            real(c_float), intent(out) :: out(*)
        end function

        !! kfs_top, kfs_bot, kob_top, kob_bot, kob over (i,j) of the memory box
        integer(c_int) function swpc3d_set_surface(ctx, kfs_top, kfs_bot, kob_top, kob_bot, kob) &
            bind(C, name='swpc3d_set_surface')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), intent(in) :: kfs_top(*), kfs_bot(*), kob_top(*), kob_bot(*), kob(*)
        end function

        !! damping profiles g(4, 1:n) at centres and edges (m_absorb_p.f90:75-93)
        integer(c_int) function swpc3d_set_pml(ctx, gxc, gxe, gyc, gye, gzc, gze) bind(C, name='swpc3d_set_pml')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            real(c_float), intent(in) :: gxc(*), gxe(*), gyc(*), gye(*), gzc(*), gze(*)
        end function

        !! m6(6, nsrc); sdrop(nsrc, nt) in FP64 (m_source.f90:775-833)
        integer(c_int) function swpc3d_set_sources(ctx, nsrc, isrc, jsrc, ksrc, m6, nt, sdrop) &
            bind(C, name='swpc3d_set_sources')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsrc, nt
            integer(c_int), intent(in) :: isrc(*), jsrc(*), ksrc(*)
            real(c_double), intent(in) :: m6(*), sdrop(*)
        end function

        integer(c_int) function swpc3d_set_bodyforce(ctx, nsrc, isrc, jsrc, ksrc, f3, nt, fdt) &
            bind(C, name='swpc3d_set_bodyforce')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsrc, nt
            integer(c_int), intent(in) :: isrc(*), jsrc(*), ksrc(*)
            real(c_double), intent(in) :: f3(*), fdt(*)
        end function

        integer(c_int) function swpc3d_set_receivers(ctx, nst, ist, jst, kst, ntdec_w, ntw) &
            bind(C, name='swpc3d_set_receivers')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: nst, ntdec_w, ntw
            integer(c_int), intent(in) :: ist(*), jst(*), kst(*)
        end function

        integer(c_int) function swpc3d_update_stress(ctx) bind(C, name='swpc3d_update_stress')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function swpc3d_update_vel(ctx) bind(C, name='swpc3d_update_vel')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function swpc3d_stressglut(ctx, it) bind(C, name='swpc3d_stressglut')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function

        integer(c_int) function swpc3d_bodyforce(ctx, it) bind(C, name='swpc3d_bodyforce')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function

        integer(c_int) function swpc3d_store_receivers(ctx, it) bind(C, name='swpc3d_store_receivers')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function

        !! ijk(3, npts) 1-based cells of a snapshot section (snap__write, m_snap.f90)
        integer(c_int) function swpc3d_add_snapshot(ctx, kind, npts, ijk, ntdec_s, nbuf, uc, m0, dt, id) &
            bind(C, name='swpc3d_add_snapshot')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            integer(c_int), value :: kind, npts, ntdec_s, nbuf
            integer(c_int), intent(in) :: ijk(*)
            real(c_float), value :: uc, m0, dt
            integer(c_int), intent(out) :: id
        end function

        integer(c_int) function swpc3d_snapshots(ctx, it) bind(C, name='swpc3d_snapshots')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function

        !! out(npts, nvar) of frame (0-based)
        integer(c_int) function swpc3d_get_snapshot(ctx, id, frame, out) bind(C, name='swpc3d_get_snapshot')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            integer(c_int), value :: id, frame
            real(c_float), intent(out) :: out(*)
        end function

        !! max-V/H/A of a 'v' or 'u' stream (output__put_maxval, m_snap.f90:2291-2344)
        integer(c_int) function swpc3d_snapshot_peaks(ctx, id) bind(C, name='swpc3d_snapshot_peaks')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: id
        end function

        integer(c_int) function swpc3d_finish_snapshots(ctx, it) bind(C, name='swpc3d_finish_snapshots')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function

        !! out(npts, 3)
        integer(c_int) function swpc3d_get_peaks(ctx, id, out) bind(C, name='swpc3d_get_peaks')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            integer(c_int), value :: id
            real(c_float), intent(out) :: out(*)
        end function

        integer(c_int) function swpc3d_step(ctx, it) bind(C, name='swpc3d_step')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function

        integer(c_int) function swpc3d_run(ctx, it0, nsteps) bind(C, name='swpc3d_run')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it0, nsteps
        end function

        !! vel(3, nst, ntw) in solver units (z positive up), disp(3, nst, ntw) FP64
        integer(c_int) function swpc3d_get_receivers(ctx, vel, disp) bind(C, name='swpc3d_get_receivers')
            import :: c_int, c_ptr, c_float, c_double
            type(c_ptr), value :: ctx
            real(c_float), intent(out) :: vel(*)
            real(c_double), intent(out) :: disp(*)
        end function

        integer(c_int) function swpc3d_vmax(ctx, out3) bind(C, name='swpc3d_vmax')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            real(c_float), intent(out) :: out3(3)
        end function

        integer(c_int) function swpc3d_sync(ctx) bind(C, name='swpc3d_sync')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        ! ---------------------------------------------------------------- P-SV

        integer(c_int) function swpc2d_create(desc, ctx) bind(C, name='swpc2d_create')
            import :: c_int, c_ptr, swpc2d_desc
            type(swpc2d_desc), intent(in) :: desc
            type(c_ptr), intent(out) :: ctx
        end function

        integer(c_int) function swpc2d_destroy(ctx) bind(C, name='swpc2d_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function swpc2d_set_medium(ctx, rho, lam, mu, taup, taus) bind(C, name='swpc2d_set_medium')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            real(c_float), intent(in) :: rho(*), lam(*), mu(*), taup(*), taus(*)
        end function

        integer(c_int) function swpc2d_set_surface(ctx, kfs_top, kfs_bot, kob_top, kob_bot, kob) &
            bind(C, name='swpc2d_set_surface')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), intent(in) :: kfs_top(*), kfs_bot(*), kob_top(*), kob_bot(*), kob(*)
        end function

        integer(c_int) function swpc2d_set_pml(ctx, gxc, gxe, gzc, gze) bind(C, name='swpc2d_set_pml')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            real(c_float), intent(in) :: gxc(*), gxe(*), gzc(*), gze(*)
        end function

        integer(c_int) function swpc2d_set_sources(ctx, nsrc, isrc, ksrc, m3, nt, sdrop) &
            bind(C, name='swpc2d_set_sources')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsrc, nt
            integer(c_int), intent(in) :: isrc(*), ksrc(*)
            real(c_double), intent(in) :: m3(*), sdrop(*)
        end function

        !! f2(2, nsrc) = (fx, fz)/M0; fdt(nsrc, nt) (source__bodyforce, src/swpc_psv/m_source.f90:616-654)
        integer(c_int) function swpc2d_set_bodyforce(ctx, nsrc, isrc, ksrc, f2, nt, fdt) &
            bind(C, name='swpc2d_set_bodyforce')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsrc, nt
            integer(c_int), intent(in) :: isrc(*), ksrc(*)
            real(c_double), intent(in) :: f2(*), fdt(*)
        end function

        integer(c_int) function swpc2d_set_receivers(ctx, nst, ist, kst, ntdec_w, ntw) &
            bind(C, name='swpc2d_set_receivers')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: nst, ntdec_w, ntw
            integer(c_int), intent(in) :: ist(*), kst(*)
        end function

        integer(c_int) function swpc2d_update_stress(ctx) bind(C, name='swpc2d_update_stress')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function swpc2d_update_vel(ctx) bind(C, name='swpc2d_update_vel')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function swpc2d_stressglut(ctx, it) bind(C, name='swpc2d_stressglut')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function

        integer(c_int) function swpc2d_bodyforce(ctx, it) bind(C, name='swpc2d_bodyforce')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function

        integer(c_int) function swpc2d_store_receivers(ctx, it) bind(C, name='swpc2d_store_receivers')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function

        !! ik(2, npts) 1-based cells of the xz section (snap__write, src/swpc_psv/m_snap.f90)
        integer(c_int) function swpc2d_add_snapshot(ctx, kind, npts, ik, ntdec_s, nbuf, uc, m0, dt, id) &
            bind(C, name='swpc2d_add_snapshot')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            integer(c_int), value :: kind, npts, ntdec_s, nbuf
            integer(c_int), intent(in) :: ik(*)
            real(c_float), value :: uc, m0, dt
            integer(c_int), intent(out) :: id
        end function

        integer(c_int) function swpc2d_snapshots(ctx, it) bind(C, name='swpc2d_snapshots')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function

        !! frame 0, 1, ... of stream id: out(npts, 2)
        integer(c_int) function swpc2d_get_snapshot(ctx, id, frame, out) bind(C, name='swpc2d_get_snapshot')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            integer(c_int), value :: id, frame
            real(c_float), intent(out) :: out(*)
        end function

        integer(c_int) function swpc2d_set_graph(ctx, steps) bind(C, name='swpc2d_set_graph')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: steps
        end function

        integer(c_int) function swpc2d_set_fusion(ctx, on) bind(C, name='swpc2d_set_fusion')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: on
        end function

        integer(c_int) function swpc2d_step(ctx, it) bind(C, name='swpc2d_step')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function

        integer(c_int) function swpc2d_run(ctx, it0, nsteps) bind(C, name='swpc2d_run')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it0, nsteps
        end function

        integer(c_int) function swpc2d_get_receivers(ctx, vel, disp) bind(C, name='swpc2d_get_receivers')
            import :: c_int, c_ptr, c_float, c_double
            type(c_ptr), value :: ctx
            real(c_float), intent(out) :: vel(*)
            real(c_double), intent(out) :: disp(*)
        end function

        integer(c_int) function swpc2d_vmax(ctx, out2) bind(C, name='swpc2d_vmax')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            real(c_float), intent(out) :: out2(2)
        end function

        integer(c_int) function swpc2d_sync(ctx) bind(C, name='swpc2d_sync')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        ! ---------------------------------------- SH (swpc_sh)

        integer(c_int) function swpcsh_create(desc, ctx) bind(C, name='swpcsh_create')
            import :: c_int, c_ptr, swpc2d_desc
            type(swpc2d_desc), intent(in) :: desc
            type(c_ptr), intent(out) :: ctx
        end function

        integer(c_int) function swpcsh_destroy(ctx) bind(C, name='swpcsh_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function swpcsh_set_medium(ctx, rho, mu, taus) bind(C, name='swpcsh_set_medium')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            real(c_float), intent(in) :: rho(*), mu(*), taus(*)
        end function

        integer(c_int) function swpcsh_set_surface(ctx, kfs_top, kfs_bot, kob_top, kob_bot, kob) &
            bind(C, name='swpcsh_set_surface')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), intent(in) :: kfs_top(*), kfs_bot(*), kob_top(*), kob_bot(*), kob(*)
        end function

        integer(c_int) function swpcsh_set_pml(ctx, gxc, gxe, gzc, gze) bind(C, name='swpcsh_set_pml')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            real(c_float), intent(in) :: gxc(*), gxe(*), gzc(*), gze(*)
        end function

        integer(c_int) function swpcsh_set_sources(ctx, nsrc, isrc, ksrc, m2, nt, sdrop) &
            bind(C, name='swpcsh_set_sources')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsrc, nt
            integer(c_int), intent(in) :: isrc(*), ksrc(*)
            real(c_double), intent(in) :: m2(*), sdrop(*)
        end function

        !! fy(nsrc) = fy/M0; fdt(nsrc, nt) (source__bodyforce, src/swpc_sh/m_source.f90:596-629)
        integer(c_int) function swpcsh_set_bodyforce(ctx, nsrc, isrc, ksrc, fy, nt, fdt) &
            bind(C, name='swpcsh_set_bodyforce')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsrc, nt
            integer(c_int), intent(in) :: isrc(*), ksrc(*)
            real(c_double), intent(in) :: fy(*), fdt(*)
        end function

        integer(c_int) function swpcsh_set_receivers(ctx, nst, ist, kst, ntdec_w, ntw) &
            bind(C, name='swpcsh_set_receivers')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: nst, ntdec_w, ntw
            integer(c_int), intent(in) :: ist(*), kst(*)
        end function

        integer(c_int) function swpcsh_update_stress(ctx) bind(C, name='swpcsh_update_stress')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function swpcsh_update_vel(ctx) bind(C, name='swpcsh_update_vel')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function swpcsh_stressglut(ctx, it) bind(C, name='swpcsh_stressglut')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function

        integer(c_int) function swpcsh_bodyforce(ctx, it) bind(C, name='swpcsh_bodyforce')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function

        integer(c_int) function swpcsh_store_receivers(ctx, it) bind(C, name='swpcsh_store_receivers')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function

        !! ik(2, npts) 1-based cells; kind SWPC_SNAP_V (Vy) or SWPC_SNAP_U (Uy) (src/swpc_sh/m_snap.f90)
        integer(c_int) function swpcsh_add_snapshot(ctx, kind, npts, ik, ntdec_s, nbuf, uc, m0, dt, id) &
            bind(C, name='swpcsh_add_snapshot')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            integer(c_int), value :: kind, npts, ntdec_s, nbuf
            integer(c_int), intent(in) :: ik(*)
            real(c_float), value :: uc, m0, dt
            integer(c_int), intent(out) :: id
        end function

        integer(c_int) function swpcsh_snapshots(ctx, it) bind(C, name='swpcsh_snapshots')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function

        !! frame 0, 1, ... of stream id: out(npts)
        integer(c_int) function swpcsh_get_snapshot(ctx, id, frame, out) bind(C, name='swpcsh_get_snapshot')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            integer(c_int), value :: id, frame
            real(c_float), intent(out) :: out(*)
        end function

        integer(c_int) function swpcsh_set_graph(ctx, steps) bind(C, name='swpcsh_set_graph')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: steps
        end function

        integer(c_int) function swpcsh_step(ctx, it) bind(C, name='swpcsh_step')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it
        end function

        integer(c_int) function swpcsh_run(ctx, it0, nsteps) bind(C, name='swpcsh_run')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: it0, nsteps
        end function

        integer(c_int) function swpcsh_get_receivers(ctx, vel, disp) bind(C, name='swpcsh_get_receivers')
            import :: c_int, c_ptr, c_float, c_double
            type(c_ptr), value :: ctx
            real(c_float), intent(out) :: vel(*)
            real(c_double), intent(out) :: disp(*)
        end function

        integer(c_int) function swpcsh_vmax(ctx, out1) bind(C, name='swpcsh_vmax')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            real(c_float), intent(out) :: out1(1)
        end function

        integer(c_int) function swpcsh_sync(ctx) bind(C, name='swpcsh_sync')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        ! ------------------------------------------ learned absorbing boundary
        ! (no counterpart in the reference: replaces absorb__update_* on a
        ! grid created with na = 0; params(nparam, 10) as in swpc_hip.h)

        integer(c_int) function swpc_lab_nparam() bind(C, name='swpc_lab_nparam')
            import :: c_int
        end function

        integer(c_int) function swpc3d_set_lab(ctx, params) bind(C, name='swpc3d_set_lab')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            real(c_float), intent(in) :: params(*)
        end function

        integer(c_int) function swpc3d_set_lab_sponge(ctx, sponge) bind(C, name='swpc3d_set_lab_sponge')
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: ctx
            real(c_float), value :: sponge
        end function

        ! ---------------------------------------- batches of independent runs

        integer(c_int) function swpc_batch_create(b) bind(C, name='swpc_batch_create')
            import :: c_int, c_ptr
            type(c_ptr), intent(out) :: b
        end function

        integer(c_int) function swpc_batch_destroy(b) bind(C, name='swpc_batch_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: b
        end function

        integer(c_int) function swpc_batch_add3d(b, ctx) bind(C, name='swpc_batch_add3d')
            import :: c_int, c_ptr
            type(c_ptr), value :: b, ctx
        end function

        integer(c_int) function swpc_batch_add2d(b, ctx) bind(C, name='swpc_batch_add2d')
            import :: c_int, c_ptr
            type(c_ptr), value :: b, ctx
        end function

        integer(c_int) function swpc_batch_run(b, it0, nsteps, graph_k) bind(C, name='swpc_batch_run')
            import :: c_int, c_ptr
            type(c_ptr), value :: b
            integer(c_int), value :: it0, nsteps, graph_k
        end function

        integer(c_int) function swpc_batch_set_grouping(b, mode) bind(C, name='swpc_batch_set_grouping')
            import :: c_int, c_ptr
            type(c_ptr), value :: b
            integer(c_int), value :: mode
        end function

        integer(c_int) function swpc_batch_sync(b) bind(C, name='swpc_batch_sync')
            import :: c_int, c_ptr
            type(c_ptr), value :: b
        end function

    end interface

contains

    !! Text of the calling thread's last error.
    function swpc_errmsg() result(msg)
        character(len=:), allocatable :: msg
        character(kind=c_char), pointer :: s(:)
        type(c_ptr) :: p
        integer :: n
        interface
            function c_strlen(str) result(n) bind(C, name='strlen')
                import :: c_ptr, c_size_t
                type(c_ptr), value :: str
                integer(c_size_t) :: n
            end function
        end interface
        p = swpc_last_error()
        n = int(c_strlen(p))
        call c_f_pointer(p, s, [n])
        allocate (character(len=n) :: msg)
        msg = transfer(s(1:n), msg)
    end function swpc_errmsg

end module swpc_hip
