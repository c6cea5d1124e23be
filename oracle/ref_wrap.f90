! ref_wrap.f90 -- TEST INFRASTRUCTURE ONLY.
!
! bind(C) entry points over the reference's own shared Fortran modules
! (src/shared/m_fdtool.f90, m_sac.f90, m_readini.f90), compiled unmodified
! from /root/reference by oracle/Makefile into oracle/_ref/libswpcref.so.
! The tests call these to pin the host-side restatements (source time
! functions, moment tensors, viscoelastic constants, index maps, SAC
! headers, input.inf parsing) against the reference's own code.
module ref_wrap

    use iso_c_binding
    use m_std
    use m_fdtool
    use m_sac
    use m_readini
    implicit none

contains

    subroutine ref_sdr2moment(strike, dip, rake, m6) bind(C, name='ref_sdr2moment')
        real(c_float), value :: strike, dip, rake
        real(c_float), intent(out) :: m6(6)   ! mxx myy mzz myz mxz mxy
        call sdr2moment(strike, dip, rake, m6(1), m6(2), m6(3), m6(4), m6(5), m6(6))
    end subroutine

    real(c_float) function ref_momentrate(t, kind, ts, tr) bind(C, name='ref_momentrate')
        real(c_float), value :: t, ts, tr
        integer(c_int), value :: kind
        real(SP) :: prm(2)
        character(16) :: name
        prm(1) = ts
        prm(2) = tr
        select case (kind)
        case (0); name = 'boxcar'
        case (1); name = 'triangle'
        case (2); name = 'herrmann'
        case (3); name = 'kupper'
        case (4); name = 'cosine'
        case (5); name = 'texp'
        case default; name = 'kupper'
        end select
        ref_momentrate = momentrate(t, trim(name), 2, prm)
    end function

    real(c_float) function ref_ikupper(t, ts, tr) bind(C, name='ref_ikupper')
        real(c_float), value :: t, ts, tr
        ref_ikupper = ikupper(t, ts, tr)
    end function

    real(c_float) function ref_seismic_moment(mw) bind(C, name='ref_seismic_moment')
        real(c_float), value :: mw
        ref_seismic_moment = seismic_moment(mw)
    end function

    subroutine ref_visco_set_relaxtime(nm, fmin, fmax, ts) bind(C, name='ref_visco_set_relaxtime')
        integer(c_int), value :: nm
        real(c_float), value :: fmin, fmax
        real(c_float), intent(out) :: ts(nm)
        call visco_set_relaxtime(nm, ts, fmin, fmax)
    end subroutine

    real(c_float) function ref_visco_constq_zeta(nm, fmin, fmax, ts) bind(C, name='ref_visco_constq_zeta')
        integer(c_int), value :: nm
        real(c_float), value :: fmin, fmax
        real(c_float), intent(in) :: ts(nm)
        ref_visco_constq_zeta = visco_constq_zeta(nm, fmin, fmax, ts)
    end function

    real(c_float) function ref_visco_chi(nm, ts, tau, f) bind(C, name='ref_visco_chi')
        integer(c_int), value :: nm
        real(c_float), intent(in) :: ts(nm)
        real(c_float), value :: tau, f
        ref_visco_chi = visco_chi(nm, ts, tau, f)
    end function

    integer(c_int) function ref_x2i(x, xbeg, dx) bind(C, name='ref_x2i')
        real(c_float), value :: x, xbeg, dx
        ref_x2i = x2i(x, xbeg, dx)
    end function

    real(c_float) function ref_i2x(i, xbeg, dx) bind(C, name='ref_i2x')
        integer(c_int), value :: i
        real(c_float), value :: xbeg, dx
        ref_i2x = i2x(i, xbeg, dx)
    end function

    subroutine ref_vcheck(vp, vs, rho, xi, vmin, vmax, rhomin, flags) bind(C, name='ref_vcheck')
        real(c_float), intent(inout) :: vp, vs, rho
        real(c_float), value :: xi, vmin, vmax, rhomin
        integer(c_int), intent(out) :: flags(3)
        logical :: a, b, c
        a = .false.; b = .false.; c = .false.
        call vcheck(vp, vs, rho, xi, vmin, vmax, rhomin, a, b, c)
        flags(1) = merge(1, 0, a)
        flags(2) = merge(1, 0, b)
        flags(3) = merge(1, 0, c)
    end subroutine

    ! Write a SAC file through the reference writer.  fh(1:11) =
    ! (delta, b, e, depmin, depmax, stlo, stla, evlo, evla, evdp, user0),
    ! ih(1:4) = (npts, idep, nzyear, nzjday), names: kstnm, kcmpnm (8 chars).
    subroutine ref_wsac(fn, nfn, fh, ih, kstnm, kcmpnm, npts, dat) bind(C, name='ref_wsac')
        integer(c_int), value :: nfn, npts
        character(kind=c_char), intent(in) :: fn(nfn), kstnm(8), kcmpnm(8)
        real(c_double), intent(in) :: fh(11)
        integer(c_int), intent(in) :: ih(4)
        real(c_float), intent(in) :: dat(npts)
        type(sac__hdr) :: sh
        character(len=nfn) :: name
        character(8) :: a, b
        integer :: i
        do i = 1, nfn
            name(i:i) = fn(i)
        end do
        do i = 1, 8
            a(i:i) = kstnm(i)
            b(i:i) = kcmpnm(i)
        end do
        call sac__init(sh)
        sh%delta = fh(1); sh%b = fh(2); sh%e = fh(3); sh%depmin = fh(4); sh%depmax = fh(5)
        sh%stlo = fh(6); sh%stla = fh(7); sh%evlo = fh(8); sh%evla = fh(9); sh%evdp = fh(10)
        sh%user0 = fh(11)
        sh%npts = ih(1); sh%idep = ih(2); sh%nzyear = ih(3); sh%nzjday = ih(4)
        sh%kstnm = a
        sh%kcmpnm = b
        call sac__write(name, sh, dat(1:npts), .true.)
    end subroutine

    ! readini over a file: kind 0 = integer, 1 = real, 2 = double, 3 = string, 4 = logical.
    subroutine ref_readini(fn, nfn, key, nkey, kind, ival, rval, dval, sval, lval) bind(C, name='ref_readini')
        integer(c_int), value :: nfn, nkey, kind
        character(kind=c_char), intent(in) :: fn(nfn), key(nkey)
        integer(c_int), intent(inout) :: ival
        real(c_float), intent(inout) :: rval
        real(c_double), intent(inout) :: dval
        character(kind=c_char), intent(inout) :: sval(256)
        integer(c_int), intent(inout) :: lval
        character(len=nfn) :: name
        character(len=nkey) :: k
        character(256) :: s, sdef
        integer :: io, i, idef
        real(SP) :: rdef
        real(DP) :: ddef
        logical :: l, ldef
        do i = 1, nfn
            name(i:i) = fn(i)
        end do
        do i = 1, nkey
            k(i:i) = key(i)
        end do
        idef = ival; rdef = rval; ddef = dval; ldef = (lval /= 0)
        sdef = ''
        do i = 1, 256
            if (sval(i) == c_null_char) exit
            sdef(i:i) = sval(i)
        end do
        open (newunit=io, file=name, action='read', status='old')
        select case (kind)
        case (0); call readini(io, k, ival, idef)
        case (1); call readini(io, k, rval, rdef)
        case (2); call readini(io, k, dval, ddef)
        case (3)
            call readini(io, k, s, trim(sdef))
            do i = 1, 255
                sval(i) = s(i:i)
            end do
            sval(256) = c_null_char
        case (4)
            call readini(io, k, l, ldef)
            lval = merge(1, 0, l)
        end select
        close (io)
    end subroutine

end module ref_wrap
