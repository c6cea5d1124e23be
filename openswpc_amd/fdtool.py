"""Host-side helpers of the reference's ``m_fdtool`` (src/shared/m_fdtool.f90).

Source time functions, moment-tensor construction, viscoelastic constants
and cell/coordinate maps.  They run once at setup (or once per step on the
host, for the source time function) and follow the reference's single
precision where it stores results in ``real(SP)``: every function returns
the value the reference would hold, rounded to float32 at the same points.
"""
from __future__ import annotations

import math

import numpy as np

f32 = np.float32
PI = math.pi  # m_std.f90:13 (real(DP))


def _f(x) -> np.float32:
    return np.float32(x)


# ---------------------------------------------------------------------------
# source time functions (m_fdtool.f90:338-497)
# ---------------------------------------------------------------------------
def kupper(t, ts, tr):
    """Single-lobed Kupper moment rate (m_fdtool.f90:338-353)."""
    t, ts, tr = _f(t), _f(ts), _f(tr)
    if ts <= t <= ts + tr:
        return _f(3 * PI * math.sin(PI * float(t - ts) / float(tr)) ** 3 / (4 * float(tr)))
    return f32(0.0)


def boxcar(t, ts, tr):
    """m_fdtool.f90 boxcar: 1/tr on [ts, ts+tr]."""
    t, ts, tr = _f(t), _f(ts), _f(tr)
    return _f(f32(1.0) / tr) if ts <= t <= ts + tr else f32(0.0)


def triangle(t, ts, tr):
    """m_fdtool.f90 triangle: unit-area triangle on [ts, ts+tr]."""
    t, ts, tr = _f(t), _f(ts), _f(tr)
    if ts <= t <= ts + tr / 2:
        return _f(4 * (t - ts) / (tr * tr))
    if ts + tr / 2 < t <= ts + tr:
        return _f(-4 * (t - ts - tr) / (tr * tr))
    return f32(0.0)


def herrmann(t, ts, tr):
    """m_fdtool.f90 herrmann: piecewise-quadratic unit-area pulse."""
    t, ts, tr = _f(t), _f(ts), _f(tr)
    t1 = ts + tr / 4
    t2 = ts + 3 * tr / 4
    if ts <= t < t1:
        return _f(16 * (t - ts) ** 2 / tr ** 3)
    if t1 <= t < t2:
        return _f(-2 * (8 * (t * t + tr * ts + ts * ts - t * tr - 2 * t * ts) + tr * tr) / tr ** 3)
    if t2 <= t <= ts + tr:
        return _f(16 * (ts + tr - t) ** 2 / tr ** 3)
    return f32(0.0)


def cosine(t, ts, tr):
    """m_fdtool.f90 cosine: (1 - cos(2 pi (t-ts)/tr))/tr on [ts, ts+tr]."""
    t, ts, tr = _f(t), _f(ts), _f(tr)
    if ts <= t <= ts + tr:
        return _f((1 - math.cos(2 * PI * float(t - ts) / float(tr))) / float(tr))
    return f32(0.0)


def texp(t, ts, tr):
    """m_fdtool.f90 texp: (2 pi)^2 t/tr^2 exp(-2 pi t/tr) for t >= ts."""
    t, ts, tr = _f(t), _f(ts), _f(tr)
    if ts <= t:
        tt = float(t - ts)
        return _f((2 * PI) ** 2 * tt / float(tr * tr) * math.exp(-2 * PI * tt / float(tr)))
    return f32(0.0)


STF = {"boxcar": boxcar, "triangle": triangle, "herrmann": herrmann,
       "kupper": kupper, "cosine": cosine, "texp": texp}
STF_KIND = {"boxcar": 0, "triangle": 1, "herrmann": 2, "kupper": 3, "cosine": 4, "texp": 5}


def momentrate(t, stftype, ts, tr):
    """m_fdtool.f90:471-497; unknown names fall back to Kupper."""
    return STF.get(stftype.strip(), kupper)(t, ts, tr)


def ikupper(t, ts, tr):
    """Time-integrated Kupper function (m_fdtool.f90:578-597); tends to 1."""
    t, ts, tr = _f(t), _f(ts), _f(tr)
    if t < ts:
        return f32(0.0)
    if t <= ts + tr:
        x = float(t - ts)
        return _f((2 + math.cos(PI * x / float(tr))) * math.sin(PI * x / (2 * float(tr))) ** 4)
    return f32(1.0)


def moment_magnitude(m0):
    """m_fdtool.f90:281-293: Mw = (log10 M0 - 9.1) * 2/3 in real(SP), -12345
    for M0 below epsilon."""
    m0 = _f(m0)
    if m0 < _f(np.finfo(np.float32).eps):
        return _f(-12345.0)
    return _f(_f(_f(np.log10(m0)) - _f(9.1)) * _f(2.0) / _f(3.0))


def seismic_moment(mw):
    """m_fdtool.f90:296-304: M0 = 10^(1.5 Mw + 9.05)."""
    return _f(10.0 ** float(_f(1.5) * _f(mw) + _f(9.05)))


# ---------------------------------------------------------------------------
# mechanism (m_fdtool.f90:307-336)
# ---------------------------------------------------------------------------
def _d2r(deg):
    # m_std.f90:132-139: real(PI/180 * deg) with PI in DP
    return _f(PI / 180.0 * float(_f(deg)))


def sdr2moment(strike, dip, rake):
    """Double-couple (strike, dip, rake) [deg] -> (mxx, myy, mzz, myz, mxz, mxy).

    Aki & Richards Box 4.4 in the x = north, y = east, z = down frame of
    the reference (m_fdtool.f90:307-336).
    """
    sind = _f(math.sin(_d2r(dip)))
    cosd = _f(math.cos(_d2r(dip)))
    sin2d = _f(math.sin(_d2r(2 * _f(dip))))
    cos2d = _f(math.cos(_d2r(2 * _f(dip))))
    sinl = _f(math.sin(_d2r(rake)))
    cosl = _f(math.cos(_d2r(rake)))
    sinf = _f(math.sin(_d2r(strike)))
    cosf = _f(math.cos(_d2r(strike)))
    sin2f = _f(math.sin(_d2r(2 * _f(strike))))
    cos2f = _f(math.cos(_d2r(2 * _f(strike))))
    mxx = -(sind * cosl * sin2f + sin2d * sinl * sinf * sinf)
    mxy = sind * cosl * cos2f + sin2d * sinl * sin2f / 2
    mxz = -(cosd * cosl * cosf + cos2d * sinl * sinf)
    myy = sind * cosl * sin2f - sin2d * sinl * cosf * cosf
    myz = -(cosd * cosl * sinf - cos2d * sinl * cosf)
    mzz = sin2d * sinl
    return tuple(_f(v) for v in (mxx, myy, mzz, myz, mxz, mxy))


# ---------------------------------------------------------------------------
# cell / coordinate maps (m_fdtool.f90:600-675)
# ---------------------------------------------------------------------------
def x2i(x, xbeg, dx):
    """ceiling((x - xbeg)/dx) in single precision."""
    return int(math.ceil(float((_f(x) - _f(xbeg)) / _f(dx))))


def i2x(i, xbeg, dx):
    """xbeg + (i - 0.5) dx in single precision (cell centre)."""
    return _f(_f(xbeg) + (_f(i) - _f(0.5)) * _f(dx))


def centres(n0, n1, beg, d):
    """i2x over i = n0..n1 as a float32 array."""
    i = np.arange(n0, n1 + 1, dtype=np.float32)
    return (np.float32(beg) + (i - np.float32(0.5)) * np.float32(d)).astype(np.float32)


# ---------------------------------------------------------------------------
# viscoelasticity (m_fdtool.f90:691-812)
# ---------------------------------------------------------------------------
def visco_set_relaxtime(nm, fmin, fmax):
    """Relaxation times log-spaced between 1/(2 pi fmax) and 1/(2 pi fmin)."""
    oa = _f(2 * PI * float(_f(fmin)))
    ob = _f(2 * PI * float(_f(fmax)))
    ts = np.zeros(nm, dtype=np.float32)
    if nm == 1:
        ts[0] = f32(1.0) / _f(math.sqrt(float(oa * ob)))
    elif nm > 1:
        for im in range(nm):
            om = _f(float(oa) * float(ob / oa) ** (im / (nm - 1)))
            ts[im] = f32(1.0) / om
    return ts


def visco_constq_zeta(nm, fmin, fmax, ts):
    """zeta of the tau-method for constant Q over [fmin, fmax] (m_fdtool.f90:756-812)."""
    if nm == 0:
        return f32(0.0)
    ts = np.asarray(ts, dtype=np.float32)
    oa = _f(2 * PI * float(_f(fmin)))
    ob = _f(2 * PI * float(_f(fmax)))
    one = f32(1.0)
    i0 = np.zeros(nm, np.float32)
    i1 = np.zeros(nm, np.float32)
    for m in range(nm):
        t = ts[m]
        i0[m] = (np.log(one + ob * ob * t * t) - np.log(one + oa * oa * t * t)) / (2 * t)
        i1[m] = ((np.arctan(ob * t) - ob * t / (1 + ob * ob * t * t))
                 - (np.arctan(oa * t) - oa * t / (1 + oa * oa * t * t))) / (2 * t)
    i2sum = f32(0.0)
    for m in range(nm - 1):
        for k in range(m + 1, nm):
            wk1 = np.arctan(ob * ts[m]) / ts[m] - np.arctan(ob * ts[k]) / ts[k]
            wk2 = np.arctan(oa * ts[m]) / ts[m] - np.arctan(oa * ts[k]) / ts[k]
            i2sum = f32(i2sum + ts[m] * ts[k] / (ts[k] ** 2 - ts[m] ** 2) * (wk1 - wk2))
    i0sum = f32(np.sum(i0, dtype=np.float32))
    i1sum = f32(np.sum(i1, dtype=np.float32))
    return f32(i0sum / (i1sum + 2 * i2sum))


def visco_chi(nm, ts, tau, f):
    """chi of the tau-method at frequency f (m_fdtool.f90:730-753), used by P-SV."""
    om = float(_f(2 * PI * float(_f(f))))
    cc = 0j
    for m in range(nm):
        w = om * float(ts[m])
        cc += complex(np.complex64((1 - 1j * w * (1 + float(tau))) / (1 - 1j * w)))
    cc = complex(np.complex64(cc / nm))
    return _f(1.0 / (cc ** -0.5).real)


def relaxed_cc(nm, ts, fq_ref):
    """Mean of i w ts/(1 - i w ts) at fq_ref, as in relaxed_medium (m_medium.f90:248-253)."""
    om = float(_f(2 * PI * float(_f(fq_ref))))
    cc = 0j
    for m in range(nm):
        w = om * float(ts[m])
        cc += complex(np.complex64((1j * w) / (1.0 - 1j * w)))
    return complex(np.complex64(cc / nm))


def vcheck(vp, vs, rho, xi, vmin, vmax, rhomin):
    """Velocity/density clamp of random media (m_fdtool.f90:15-48), vectorised.

    Returns (vp, vs, rho, flags) with flags = (vmin_under, vmax_over, rhomin_under).
    """
    vp = np.array(vp, dtype=np.float32, copy=True)
    vs = np.array(vs, dtype=np.float32, copy=True)
    rho = np.array(rho, dtype=np.float32, copy=True)
    xi = np.asarray(xi, dtype=np.float32)
    vmin, vmax, rhomin = f32(vmin), f32(vmax), f32(rhomin)
    with np.errstate(divide="ignore", invalid="ignore"):
        gamma = (vp / vs).astype(np.float32)
    gamma = np.where(gamma < np.finfo(np.float32).eps, np.float32(math.sqrt(3.0)), gamma).astype(np.float32)
    over = (vp > vmax) | (vs > vmax)
    xi2 = ((1 + xi) * vmax / vp - 1).astype(np.float32)
    rho = np.where(over, rho * (1 + f32(0.8) * xi2) / (1 + f32(0.8) * xi), rho).astype(np.float32)
    vp = np.where(over, vmax, vp).astype(np.float32)
    vs = np.where(over, vmax / gamma, vs).astype(np.float32)
    under = (vp < vmin) | (vs < vmin)
    vs = np.where(under, vmin, vs).astype(np.float32)
    vp = np.where(under, vmin * gamma, vp).astype(np.float32)
    low = rho < rhomin
    rho = np.where(low, rhomin, rho).astype(np.float32)
    return vp, vs, rho, (bool(under.any()), bool(over.any()), bool(low.any()))
