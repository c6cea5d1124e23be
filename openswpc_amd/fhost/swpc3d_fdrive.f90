!! Fortran host for the 3-D time loop, in the reference's order
!! (src/swpc_3d/main.f90:125-144), over the C ABI.
!!
!!   swpc3d_fdrive check          binding checks that need no GPU
!!   swpc3d_fdrive <in> <out>     run a case written by openswpc_amd.fcase
!!
!! The case file (stream, little-endian) holds: int32 nx ny nz na nt
!! precision nsrc nst ntdec_w ntw; float64 dx dy dz dt; float32 ts(3);
!! ten float32 medium boxes (k,i,j); five int32 (i,j) maps; six float32
!! profiles g(4,n); sources isrc jsrc ksrc (int32), m6(6,nsrc), sdrop(nsrc,nt)
!! (float64); receivers ist jst kst (int32).  The output holds
!! vel(3,nst,ntw) float32, disp(3,nst,ntw) float64 and vmax(3) float32.
program swpc3d_fdrive

    use iso_c_binding
    use iso_fortran_env, only: error_unit
    use swpc_hip
    use m_kernel_hip
    implicit none

    character(len=512) :: fin, fout
    integer :: io, it
    integer(c_int) :: hd(10), ierr
    type(swpc3d_desc) :: d
    real(c_float), allocatable :: med(:, :), gxc(:), gxe(:), gyc(:), gye(:), gzc(:), gze(:)
    integer(c_int), allocatable :: maps(:, :), isrc(:), jsrc(:), ksrc(:), ist(:), jst(:), kst(:)
    real(c_double), allocatable :: m6(:), sdrop(:), disp(:)
    real(c_float), allocatable :: vel(:)
    real(c_float) :: vm(3)
    real(c_double) :: h(4)
    integer(c_int64_t) :: nbox, nmap

    call get_command_argument(1, fin)
    if (trim(fin) == 'check') then
        call binding_checks()
        stop
    end if
    call get_command_argument(2, fout)

    open (newunit=io, file=trim(fin), access='stream', form='unformatted', action='read', status='old')
    read (io) hd
    read (io) h
    d%nx = hd(1); d%ny = hd(2); d%nz = hd(3); d%na = hd(4); d%nm = 3; d%precision = hd(6)
    d%dx = h(1); d%dy = h(2); d%dz = h(3); d%dt = h(4)
    d%ts = 0
    read (io) d%ts(1:3)
    nbox = int(d%nx + 6, c_int64_t) * (d%ny + 6) * (d%nz + 6)
    nmap = int(d%nx + 6, c_int64_t) * (d%ny + 6)
    allocate (med(nbox, 10), maps(nmap, 5))
    read (io) med
    read (io) maps
    allocate (gxc(4 * d%nx), gxe(4 * d%nx), gyc(4 * d%ny), gye(4 * d%ny), gzc(4 * d%nz), gze(4 * d%nz))
    read (io) gxc, gxe, gyc, gye, gzc, gze
    allocate (isrc(hd(7)), jsrc(hd(7)), ksrc(hd(7)), m6(6 * hd(7)), sdrop(int(hd(7), 8) * hd(5)))
    read (io) isrc, jsrc, ksrc, m6, sdrop
    allocate (ist(hd(8)), jst(hd(8)), kst(hd(8)))
    read (io) ist, jst, kst
    close (io)

    call kernel_hip__setup(d, med(:, 1), med(:, 2), med(:, 3), med(:, 4), med(:, 5), med(:, 6), med(:, 7), &
                           med(:, 8), med(:, 9), med(:, 10), maps(:, 1), maps(:, 2), maps(:, 3), maps(:, 4), &
                           maps(:, 5), gxc, gxe, gyc, gye, gzc, gze)
    call kernel_hip__check(swpc3d_set_sources(kernel_hip__ctx(), hd(7), isrc, jsrc, ksrc, m6, hd(5), sdrop), &
                           'set_sources')
    call kernel_hip__check(swpc3d_set_receivers(kernel_hip__ctx(), hd(8), ist, jst, kst, hd(9), hd(10)), &
                           'set_receivers')

    do it = 1, hd(5)
        call wav__store(it)
        call snap__write(it)
        call kernel__update_stress()
        call absorb__update_stress()
        call source__stressglut(it)
        call kernel__update_vel()
        call absorb__update_vel()
    end do

    allocate (vel(3 * hd(8) * hd(10)), disp(3 * hd(8) * hd(10)))
    call wav__get(vel, disp)
    call kernel__vmax(vm(1), vm(2), vm(3))
    call kernel_hip__finish()

    open (newunit=io, file=trim(fout), access='stream', form='unformatted', action='write', status='replace')
    write (io) vel, disp, vm
    close (io)

contains

    !! Argument passing and error reporting through the bindings, without a
    !! device: descriptors the library must reject before any HIP call.
    subroutine binding_checks()
        type(swpc3d_desc) :: b
        type(swpc2d_desc) :: b2
        type(c_ptr) :: c
        character(len=:), allocatable :: msg
        if (swpc_version() < 10000) error stop 'version'
        b = swpc3d_desc(4, 64, 64, 8, 3, SWPC_FP32, 0.1d0, 0.1d0, 0.1d0, 0.01d0, 0.0)
        ierr = swpc3d_create(b, c)
        msg = swpc_errmsg()
        if (ierr /= SWPC_EINVAL .or. index(msg, 'too small') == 0) error stop 'grid check'
        b = swpc3d_desc(64, 64, 64, 8, 2, SWPC_FP32, 0.1d0, 0.1d0, 0.1d0, 0.01d0, 0.0)
        ierr = swpc3d_create(b, c)
        if (ierr /= SWPC_EINVAL .or. index(swpc_errmsg(), 'NM = 3') == 0) error stop 'nm check'
        b = swpc3d_desc(64, 64, 64, 8, 3, SWPC_FP32, 0.1d0, -0.1d0, 0.1d0, 0.01d0, 0.0)
        b%ts(1:3) = [0.1, 0.2, 0.3]
        ierr = swpc3d_create(b, c)
        if (ierr /= SWPC_EINVAL .or. index(swpc_errmsg(), 'dx/dy/dz/dt') == 0) error stop 'dy check'
        b2 = swpc2d_desc(64, 64, 40, 3, SWPC_FP32, 0.1d0, 0.1d0, 0.01d0, 0.0)
        ierr = swpc2d_create(b2, c)
        if (ierr /= SWPC_EINVAL) error stop 'psv na check'
        if (swpc3d_update_stress(c_null_ptr) == SWPC_OK) error stop 'null ctx'
        write (*, '(A)') 'bindings ok: ' // msg
    end subroutine binding_checks

end program swpc3d_fdrive
