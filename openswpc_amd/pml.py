"""ADE-CFS PML damping profiles (absorb_p__setup / damping_profile,
src/swpc_3d/m_absorb_p.f90:60-124, 507-547; P-SV src/swpc_psv/m_absorb_p.f90:56-104, 481-521).
"""
from __future__ import annotations

import math

import numpy as np

F32 = np.float32


def damping_profile(x, H, xbeg, xend, na, fcut, dt, pa=1):
    """g(1..4) at coordinate x for an absorber of thickness H (all SP).

    R0 = 10^(-(log10(na) - 1)/log10(2) - 3), d0 = -(pd+1) cp ln(R0)/(2H),
    a0 = pi fcut, b0 = 7 with pd = 1, pb = 2, cp = 6; pa = 1 in 3-D and
    P-SV, 2 in SH (src/swpc_sh/m_absorb_p.f90:393-425).
    """
    x, H, xbeg, xend, dt = F32(x), F32(H), F32(xbeg), F32(xend), F32(dt)
    R0 = F32(10.0 ** float(-(F32(math.log10(F32(na))) - 1) / F32(math.log10(2.0)) - F32(3.0)))
    d0 = F32(-(F32(1.0) / (F32(2.0) * H)) * F32(2) * F32(6.0) * F32(math.log(R0)))
    b0 = F32(7.0)
    a0 = F32(math.pi * float(F32(fcut)))
    if x <= xbeg + H:
        xx = (xbeg + H) - x
    elif x >= xend - H:
        xx = x - (xend - H)
    else:
        xx = F32(0.0)
    r = F32(abs(xx / H))
    d = F32(d0 * r)
    a = F32(a0 * (F32(1.0) - (r if pa == 1 else F32(r * r))))
    b = F32(F32(1.0) + (b0 - F32(1.0)) * r ** 2)
    h = dt / F32(2.0)
    den = F32(1.0) + h * (a + d / b)
    g = np.empty(4, F32)
    g[0] = ((F32(1.0) + h * a) / b) / den
    g[1] = (F32(-1.0) / b) / den
    g[2] = (F32(1.0) - h * (a + d / b)) / den
    g[3] = (d / b) / den
    return g


def profiles(centres, d, n, beg, end, na, fcut, dt, H=None, pa=1):
    """(gc, ge) of shape (n, 4) at cell centres i = 1..n and at centre + d/2.

    ``centres`` holds i2x over the memory box i = -2..n+3.  H defaults to
    na*real(dx) in single precision (3-D, m_absorb_p.f90:79-81).
    """
    if H is None:
        H = F32(na * F32(d))
    gc = np.zeros((n, 4), F32)
    ge = np.zeros((n, 4), F32)
    for i in range(1, n + 1):
        x = centres[i + 2]
        gc[i - 1] = damping_profile(x, H, beg, end, na, fcut, dt, pa)
        ge[i - 1] = damping_profile(F32(x + F32(d) / F32(2.0)), H, beg, end, na, fcut, dt, pa)
    return gc, ge


def pml3(g, fcut):
    """(gxc, gxe, gyc, gye, gzc, gze) for a Grid3D."""
    gx = profiles(g.xc, g.dx, g.nx, g.xbeg, g.xend, g.na, fcut, g.dt)
    gy = profiles(g.yc, g.dy, g.ny, g.ybeg, g.yend, g.na, fcut, g.dt)
    gz = profiles(g.zc, g.dz, g.nz, g.zbeg, g.zend, g.na, fcut, g.dt)
    return gx[0], gx[1], gy[0], gy[1], gz[0], gz[1]


def pml2(g, fcut):
    """(gxc, gxe, gzc, gze) for a Grid2D; the P-SV absorber uses H = real(na*dx)."""
    gx = profiles(g.xc, g.dx, g.nx, g.xbeg, g.xend, g.na, fcut, g.dt, H=F32(g.na * g.dx))
    gz = profiles(g.zc, g.dz, g.nz, g.zbeg, g.zend, g.na, fcut, g.dt, H=F32(g.na * g.dz))
    return gx[0], gx[1], gz[0], gz[1]


def pml_sh(g, fcut):
    """(gxc, gxe, gzc, gze) for the SH absorber on a Grid2D: H = real(na*dx)
    and pa = 2 (src/swpc_sh/m_absorb_p.f90:49-66, 393-425)."""
    gx = profiles(g.xc, g.dx, g.nx, g.xbeg, g.xend, g.na, fcut, g.dt, H=F32(g.na * g.dx), pa=2)
    gz = profiles(g.zc, g.dz, g.nz, g.zbeg, g.zend, g.na, fcut, g.dt, H=F32(g.na * g.dz), pa=2)
    return gx[0], gx[1], gz[0], gz[1]
