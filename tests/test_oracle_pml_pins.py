"""Pins of the oracle's ADE-CFS PML updates by their closed forms.

Under a stress (or velocity) field that is linear in x, y and z, every
2nd-order difference of the absorber update is a constant, so each ADE
value follows its recurrence a_n = g3 a_(n-1) + g4 c dt in closed form,

    a_n = g4 c dt (1 + g3 + ... + g3^(n-1)),

and each velocity (stress) increment is b (sum over directions of
g1 c + g2 a_(n-1)) dt, with a_(n-1) the value *before* the step's ADE
update (src/swpc_3d/m_absorb_p.f90:266-298, 437-494;
src/swpc_psv/m_absorb_p.f90:172-203, 345-386).  The fields get 18 (3-D)
or 8 (P-SV) distinct gradients and the buoyancies and shear moduli
distinct values, so a swapped centre/edge profile, an ADE array paired
with the wrong derivative or profile, a wrong sign, an update done in the
wrong order or a dropped term fails; cells outside the absorber must stay
untouched.

The plane-wave test measures what the whole absorber does with a normal-
incidence P pulse against the design reflection coefficient R0
(src/swpc_3d/m_absorb_p.f90:525-530).
"""
import math

import numpy as np
import pytest

from openswpc_amd import medium as md, pml, problem
from openswpc_amd.source import stress_glut2, stress_glut3
from oracle.oracle import Oracle2D, Oracle3D

NSTEP = 4


def _case3(nx=22, ny=20, nz=18, na=5, dx=0.5, dt=0.04):
    g = md.Grid3D(nx, ny, nz, dx, dx, dx, dt, na=na, zbeg=-dx * 3)
    rho, lam, mu, qp, qs = md.layered3(g, [md.Layer(-1000.0, 2.7, 6.0, 3.5, 1e30, 1e30)])
    m = md.medium_setup3(g, rho, lam, mu, qp, qs, benchmark=True)
    return problem.Case3D(g, m, pml.pml3(g, 1.0), stress_glut3(g, [], 2), np.zeros((0, 3), np.int32), [], 2, 1,
                          (0, 0, 0))


def _case2(nx=24, nz=18, na=5, dx=0.5, dt=0.04):
    g = md.Grid2D(nx, nz, dx, dx, dt, na=na, zbeg=-dx * 3)
    rho, lam, mu, qp, qs = md.layered2(g, [md.Layer(-1000.0, 2.7, 6.0, 3.5, 1e30, 1e30)])
    m = md.medium_setup2(g, rho, lam, mu, qp, qs, benchmark=True)
    return problem.Case2D(g, m, pml.pml2(g, 1.0), stress_glut2(g, [], 2), np.zeros((0, 2), np.int32), [], 2, 1,
                          (0, 0, 0))


def _box_index3(g):
    """Cell indices (J, I, K) over the (j, i, k) memory box [-2, n+3]."""
    j = np.arange(-2, g.ny + 4)[:, None, None]
    i = np.arange(-2, g.nx + 4)[None, :, None]
    k = np.arange(-2, g.nz + 4)[None, None, :]
    return j, i, k


def _prof3(g, prof, axis, q):
    """Profile column q (1..4) of gxc/gxe/... broadcast over the 3-D box;
    NaN outside 1..n so that any use of an out-of-range entry shows."""
    n = (g.nx, g.ny, g.nz)[axis]
    v = np.full(n + 6, np.nan)
    v[3:3 + n] = prof[:, q - 1].astype(np.float64)
    shp = [1, 1, 1]
    shp[(1, 0, 2)[axis]] = n + 6
    return v.reshape(shp)


def _pml_mask3(g):
    j, i, k = _box_index3(g)
    inside = (i >= 1) & (i <= g.nx) & (j >= 1) & (j <= g.ny) & (k >= 1) & (k <= g.nz)
    side = (i <= g.na) | (i >= g.nx - g.na + 1) | (j <= g.na) | (j >= g.ny - g.na + 1)
    kbeg = np.where(side, 1, g.nz - g.na + 1)     # kbeg_a, src/swpc_3d/m_global.f90:336-345
    return inside & (k >= kbeg)


def _geom_sum(g3, n):
    """1 + g3 + ... + g3^(n-1)."""
    return sum(g3 ** m for m in range(n))


def test_3d_pml_velocity_ade_closed_form():
    case = _case3()
    g = case.grid
    o = Oracle3D(case, 64)
    b = {"bx": 0.375, "by": 0.625, "bz": 0.8125}
    for n, v in b.items():
        o.med[n][...] = v
    J, I, K = _box_index3(g)
    # distinct gradients (per km) of each stress component along x, y, z
    grad = {"Sxx": (1.1, -0.7, 0.3), "Syy": (0.45, 1.9, -1.3), "Szz": (-0.2, 0.65, 2.3),
            "Syz": (0.9, -1.6, 0.55), "Sxz": (-1.2, 0.35, 1.45), "Sxy": (2.1, -0.85, 0.15)}
    for n, (cx, cy, cz) in grad.items():
        o.set_field(n, (cx * I + cy * J + cz * K) * g.dx)
    gxc, gxe, gyc, gye, gzc, gze = case.pml
    P = {nm: {q: _prof3(g, p, ax, q) for q in (1, 2, 3, 4)}
         for nm, p, ax in (("xc", gxc, 0), ("xe", gxe, 0), ("yc", gyc, 1), ("ye", gye, 1), ("zc", gzc, 2),
                           ("ze", gze, 2))}
    # ADE array -> (stress, derivative axis, profile)
    ade = {"axSxx": ("Sxx", 0, "xe"), "aySxy": ("Sxy", 1, "yc"), "azSxz": ("Sxz", 2, "zc"),
           "axSxy": ("Sxy", 0, "xc"), "aySyy": ("Syy", 1, "ye"), "azSyz": ("Syz", 2, "zc"),
           "axSxz": ("Sxz", 0, "xc"), "aySyz": ("Syz", 1, "yc"), "azSzz": ("Szz", 2, "ze")}
    vel = {"Vx": ("bx", ("axSxx", "aySxy", "azSxz")), "Vy": ("by", ("axSxy", "aySyy", "azSyz")),
           "Vz": ("bz", ("axSxz", "aySyz", "azSzz"))}
    dt = float(np.float32(g.dt))
    mask = _pml_mask3(g)
    prev = {a: 0.0 for a in ade}
    vprev = {v: np.zeros(g.shape) for v in vel}
    for step in range(1, NSTEP + 1):
        o.L.o3d_absorb_vel(o._ps())
        exp_a = {}
        for a, (sn, ax, pn) in ade.items():
            c = grad[sn][ax]
            exp_a[a] = P[pn][4] * c * dt * _geom_sum(P[pn][3], step)
        for v, (bn, arrs) in vel.items():
            inc = 0.0
            for a in arrs:
                sn, ax, pn = ade[a]
                inc = inc + P[pn][1] * grad[sn][ax] + P[pn][2] * prev[a]
            e = vprev[v] + b[bn] * inc * dt
            got = o.f[v]
            assert np.allclose(got[mask], np.broadcast_to(e, got.shape)[mask], rtol=1e-11, atol=1e-13), (v, step)
            assert np.all(got[~mask] == 0), v
            vprev[v] = np.broadcast_to(e, got.shape).copy()
        for a, e in exp_a.items():
            got = o.A[a]
            assert np.allclose(got[mask], np.broadcast_to(e, got.shape)[mask], rtol=1e-11, atol=1e-15), (a, step)
            assert np.all(got[~mask] == 0), a
        prev = exp_a
    # the damping fired: ADE values are far from zero and V depends on them
    assert max(np.abs(o.A[a][mask]).max() for a in ade) > 0.1


def test_3d_pml_stress_ade_closed_form():
    case = _case3()
    g = case.grid
    o = Oracle3D(case, 64)
    lam, mu = 3.25, 2.5           # lam + 2 mu and lam are exact in FP32
    mij = {"muyz": 1.25, "muxz": 1.75, "muxy": 2.125}
    o.med["lam"][...] = lam
    o.med["mu"][...] = mu
    for n, v in mij.items():
        o.med[n][...] = v
    J, I, K = _box_index3(g)
    grad = {"Vx": (0.8, -1.4, 0.6), "Vy": (1.7, 0.25, -0.95), "Vz": (-0.55, 1.15, 2.05)}
    for n, (cx, cy, cz) in grad.items():
        o.set_field(n, (cx * I + cy * J + cz * K) * g.dx)
    gxc, gxe, gyc, gye, gzc, gze = case.pml
    P = {nm: {q: _prof3(g, p, ax, q) for q in (1, 2, 3, 4)}
         for nm, p, ax in (("xc", gxc, 0), ("xe", gxe, 0), ("yc", gyc, 1), ("ye", gye, 1), ("zc", gzc, 2),
                           ("ze", gze, 2))}
    ade = {"axVx": ("Vx", 0, "xc"), "ayVy": ("Vy", 1, "yc"), "azVz": ("Vz", 2, "zc"),
           "ayVx": ("Vx", 1, "ye"), "azVx": ("Vx", 2, "ze"), "axVy": ("Vy", 0, "xe"),
           "azVy": ("Vy", 2, "ze"), "axVz": ("Vz", 0, "xe"), "ayVz": ("Vz", 1, "ye")}
    dt = float(np.float32(g.dt))
    mask = _pml_mask3(g)
    prev = {a: 0.0 for a in ade}
    sprev = {s: np.zeros(g.shape) for s in ("Sxx", "Syy", "Szz", "Syz", "Sxz", "Sxy")}

    def eps(a):
        vn, ax, pn = ade[a]
        return P[pn][1] * grad[vn][ax] + P[pn][2] * prev[a]

    for step in range(1, NSTEP + 1):
        o.L.o3d_absorb_stress(o._ps())
        exx, eyy, ezz = eps("axVx"), eps("ayVy"), eps("azVz")
        l2m = lam + 2 * mu
        exp_s = {"Sxx": l2m * exx + lam * (eyy + ezz), "Syy": l2m * eyy + lam * (exx + ezz),
                 "Szz": l2m * ezz + lam * (exx + eyy),
                 "Syz": mij["muyz"] * (eps("ayVz") + eps("azVy")),
                 "Sxz": mij["muxz"] * (eps("axVz") + eps("azVx")),
                 "Sxy": mij["muxy"] * (eps("axVy") + eps("ayVx"))}
        for s, inc in exp_s.items():
            e = np.broadcast_to(sprev[s] + inc * dt, o.f[s].shape)
            got = o.f[s]
            assert np.allclose(got[mask], e[mask], rtol=1e-11, atol=1e-13), (s, step)
            assert np.all(got[~mask] == 0), s
            sprev[s] = e.copy()
        exp_a = {}
        for a, (vn, ax, pn) in ade.items():
            exp_a[a] = P[pn][4] * grad[vn][ax] * dt * _geom_sum(P[pn][3], step)
            got = o.A[a]
            assert np.allclose(got[mask], np.broadcast_to(exp_a[a], got.shape)[mask], rtol=1e-11,
                               atol=1e-15), (a, step)
        prev = exp_a
    assert max(np.abs(o.A[a][mask]).max() for a in ade) > 0.1


def _box_index2(g):
    i = np.arange(-2, g.nx + 4)[:, None]
    k = np.arange(-2, g.nz + 4)[None, :]
    return i, k


def _prof2(g, prof, axis, q):
    n = (g.nx, g.nz)[axis]
    v = np.full(n + 6, np.nan)
    v[3:3 + n] = prof[:, q - 1].astype(np.float64)
    return v.reshape((n + 6, 1) if axis == 0 else (1, n + 6))


def _pml_mask2(g):
    i, k = _box_index2(g)
    inside = (i >= 1) & (i <= g.nx) & (k >= 1) & (k <= g.nz)
    kbeg = np.where((i <= g.na) | (i >= g.nx - g.na + 1), 1, g.nz - g.na + 1)
    return inside & (k >= kbeg)


def test_psv_pml_velocity_ade_closed_form():
    """P-SV: bx = 2/(rho(k,i) + rho(k,i+1)) and bz = 2/(rho(k,i) + rho(k+1,i))
    are formed in the absorber update (src/swpc_psv/m_absorb_p.f90:181-182);
    rho varies with i and k, so that the two differ from cell to cell."""
    case = _case2()
    g = case.grid
    o = Oracle2D(case, 64)
    I, K = _box_index2(g)
    rho = (2.0 + 0.25 * (I % 3) + 0.5 * (K % 2)).astype(np.float32)
    o.med["rho"][...] = rho
    r32 = rho.astype(np.float32)
    bx = np.zeros(g.shape)
    bz = np.zeros(g.shape)
    bx[:-1, :] = np.float32(2.0) / (r32[:-1, :] + r32[1:, :])
    bz[:, :-1] = np.float32(2.0) / (r32[:, :-1] + r32[:, 1:])
    grad = {"Sxx": (1.3, -0.4), "Szz": (0.35, 2.2), "Sxz": (-1.05, 0.75)}
    for n, (cx, cz) in grad.items():
        o.set_field(n, (cx * I + cz * K) * g.dx)
    gxc, gxe, gzc, gze = case.pml
    P = {nm: {q: _prof2(g, p, ax, q) for q in (1, 2, 3, 4)}
         for nm, p, ax in (("xc", gxc, 0), ("xe", gxe, 0), ("zc", gzc, 1), ("ze", gze, 1))}
    ade = {"axSxx": ("Sxx", 0, "xe"), "azSxz": ("Sxz", 1, "zc"), "axSxz": ("Sxz", 0, "xc"),
           "azSzz": ("Szz", 1, "ze")}
    vel = {"Vx": (bx, ("axSxx", "azSxz")), "Vz": (bz, ("axSxz", "azSzz"))}
    dt = float(np.float32(g.dt))
    mask = _pml_mask2(g)
    prev = {a: 0.0 for a in ade}
    vprev = {v: np.zeros(g.shape) for v in vel}
    for step in range(1, NSTEP + 1):
        o.L.o2d_absorb_vel(o._ps())
        for v, (b, arrs) in vel.items():
            inc = sum(P[ade[a][2]][1] * grad[ade[a][0]][ade[a][1]] + P[ade[a][2]][2] * prev[a] for a in arrs)
            e = np.broadcast_to(vprev[v] + b * inc * dt, g.shape)
            assert np.allclose(o.f[v][mask], e[mask], rtol=1e-11, atol=1e-13), (v, step)
            assert np.all(o.f[v][~mask] == 0), v
            vprev[v] = e.copy()
        exp_a = {}
        for a, (sn, ax, pn) in ade.items():
            exp_a[a] = P[pn][4] * grad[sn][ax] * dt * _geom_sum(P[pn][3], step)
            got = o.A[a]
            assert np.allclose(got[mask], np.broadcast_to(exp_a[a], g.shape)[mask], rtol=1e-11, atol=1e-15), (a, step)
            assert np.all(got[~mask] == 0), a
        prev = exp_a
    assert max(np.abs(o.A[a][mask]).max() for a in ade) > 0.1


def test_psv_pml_stress_ade_closed_form():
    """P-SV: mu_xz is the 4-point harmonic mean formed in the update
    (src/swpc_psv/m_absorb_p.f90:367-371); mu takes four values in a 2 x 2
    pattern, so every node's mean is 4 / sum(1/mu_q)."""
    case = _case2()
    g = case.grid
    o = Oracle2D(case, 64)
    I, K = _box_index2(g)
    lam = 3.25
    mus = np.array([[2.5, 1.25], [3.0, 1.75]])
    mu = mus[I % 2, K % 2]
    o.med["lam"][...] = lam
    o.med["mu"][...] = mu
    hm = 4.0 / (1 / 2.5 + 1 / 1.25 + 1 / 3.0 + 1 / 1.75)
    grad = {"Vx": (0.8, -1.4), "Vz": (-0.55, 2.05)}
    for n, (cx, cz) in grad.items():
        o.set_field(n, (cx * I + cz * K) * g.dx)
    gxc, gxe, gzc, gze = case.pml
    P = {nm: {q: _prof2(g, p, ax, q) for q in (1, 2, 3, 4)}
         for nm, p, ax in (("xc", gxc, 0), ("xe", gxe, 0), ("zc", gzc, 1), ("ze", gze, 1))}
    ade = {"axVx": ("Vx", 0, "xc"), "azVz": ("Vz", 1, "zc"), "azVx": ("Vx", 1, "ze"), "axVz": ("Vz", 0, "xe")}
    dt = float(np.float32(g.dt))
    mask = _pml_mask2(g)
    prev = {a: 0.0 for a in ade}
    sprev = {s: np.zeros(g.shape) for s in ("Sxx", "Szz", "Sxz")}

    def eps(a):
        vn, ax, pn = ade[a]
        return P[pn][1] * grad[vn][ax] + P[pn][2] * prev[a]

    for step in range(1, NSTEP + 1):
        o.L.o2d_absorb_stress(o._ps())
        exx, ezz = eps("axVx"), eps("azVz")
        l2m = lam + 2 * mu
        exp_s = {"Sxx": l2m * exx + lam * ezz, "Szz": l2m * ezz + lam * exx,
                 "Sxz": hm * (eps("axVz") + eps("azVx"))}
        for s, inc in exp_s.items():
            e = np.broadcast_to(sprev[s] + inc * dt, g.shape)
            tol = 1e-6 if s == "Sxz" else 1e-11     # the harmonic mean is formed in FP32 with +epsilon(1.0)
            assert np.allclose(o.f[s][mask], e[mask], rtol=tol, atol=1e-13), (s, step)
            assert np.all(o.f[s][~mask] == 0), s
            sprev[s] = e.copy()
        exp_a = {}
        for a, (vn, ax, pn) in ade.items():
            exp_a[a] = P[pn][4] * grad[vn][ax] * dt * _geom_sum(P[pn][3], step)
            assert np.allclose(o.A[a][mask], np.broadcast_to(exp_a[a], g.shape)[mask], rtol=1e-11, atol=1e-15), a
        prev = exp_a


def _plane_wave_reflection(na, w, nx=800, nint=200, dx=0.05, dt=0.004):
    """Peak reflected / peak incident Vz of a down-going P plane pulse
    (Gaussian of half-width w cells, vp = 6 km/s = the profile's assumed
    cp) at normal incidence on the bottom absorber, read at the centre
    column, where the ends of the finite wavefront have no influence
    within the window."""
    g = md.Grid2D(nx, nint + na, dx, dx, dt, na=na, zbeg=2.0)
    a, b, rho = 6.0, 3.5, 2.7
    r_, lam, mu, qp, qs = md.layered2(g, [md.Layer(-1000.0, rho, a, b, 1e30, 1e30)])
    m = md.medium_setup2(g, r_, lam, mu, qp, qs, benchmark=True)
    case = problem.Case2D(g, m, pml.pml2(g, 1.0), stress_glut2(g, [], 2), np.zeros((0, 2), np.int32), [], 2, 1,
                          (0, 0, 0))
    o = Oracle2D(case, 64)
    k = np.arange(-2, g.nz + 4).astype(float)
    z0, krec, ic = 60.0, 110, nx // 2
    ones = np.ones((g.nx + 6, 1))
    # Vz at (k + 1/2) dz and half a step ahead of the stresses
    o.set_field("Vz", np.exp(-((k + 0.5 - a * dt / (2 * dx) - z0) / w) ** 2)[None, :] * ones)
    s = -rho * a * np.exp(-((k - z0) / w) ** 2)[None, :] * ones
    o.set_field("Szz", s)
    o.set_field("Sxx", s * (1 - 2 * b * b / (a * a)))
    nsteps = int(((nint - z0) + (nint - krec) + 3 * w) * dx / a / dt) + 40
    tr = np.zeros(nsteps)
    for it in range(nsteps):
        o.update_stress()
        o.update_vel()
        tr[it] = o.f["Vz"][ic + 2, krec + 2]
    t = (np.arange(nsteps) + 1) * dt
    t_inc = (krec - z0) * dx / a
    t_ref = (2 * (nint + 0.5) - z0 - krec) * dx / a
    half = 0.5 * (t_ref - t_inc)
    return np.abs(tr[np.abs(t - t_ref) < half]).max() / np.abs(tr[np.abs(t - t_inc) < half]).max()


def test_pml_normal_incidence_reflection_falls_with_thickness():
    """A P pulse at normal incidence: the reflection falls with the absorber
    thickness, and at na = 6, where the design R0 = 10^(-(log10 na - 1) /
    log10 2 - 3) = 5.5e-3 is above the discretisation floor, it is within
    a factor 2 of R0.  A wrong damping amplitude (d0), a wrong sign of the
    ADE coupling or a mispaired profile moves it by far more."""
    R0 = {na: 10 ** (-(math.log10(na) - 1) / math.log10(2) - 3) for na in (6, 10, 20, 60)}
    R = {na: _plane_wave_reflection(na, 8.0) for na in R0}
    print({na: f"{R[na]:.2e} (R0 {R0[na]:.1e})" for na in R})
    nas = sorted(R)
    assert all(R[a] > R[b] for a, b in zip(nas, nas[1:]))
    assert R0[6] / 2 < R[6] < 2 * R0[6]
    assert R[60] < R[6] / 10


def _region_error(na, top_damping=True, nxu=120, nzu=80, dx=0.25, dt=0.01, nt=500, ref=None):
    """rms over time and over the region U (nxu x nzu cells below the top
    boundary) of the velocity difference between a grid whose absorber
    (na cells at x-, x+ and the bottom) surrounds U and an echo-free
    reflecting grid, relative to the rms of the echo-free field.  A double
    couple at 7.5 km depth, 1 s Kupper rise, uniform elastic medium.
    top_damping=False gives the z profile its undamped value in the top
    half of the grid, where the reference's profile also damps z
    (damping_profile treats zbeg like an absorbing face)."""
    from openswpc_amd import fdtool
    from openswpc_amd.source import PointSource
    ext = 0 if ref is not None else int(6.0 * nt * dt / dx / 2) + 5
    nax = na if ref is not None else 0
    nx, nz = nxu + 2 * (nax + ext), nzu + nax + ext
    g = md.Grid2D(nx, nz, dx, dx, dt, na=nax, xbeg=-(nxu / 2 + nax + ext) * dx, zbeg=0.0)
    r_, lam, mu, qp, qs = md.layered2(g, [md.Layer(-1000.0, 2.7, 6.0, 3.5, 1e30, 1e30)])
    m = md.medium_setup2(g, r_, lam, mu, qp, qs, benchmark=True)
    src = PointSource(0.3 * dx, 0.0, 30.3 * dx, 0.1, 1.0, 1e15, fdtool.sdr2moment(30.0, 60.0, 20.0))
    gl = stress_glut2(g, [src], nt)
    if nax == 0:
        prof = tuple(np.zeros((n, 4), np.float32) for n in (nx, nx, nz, nz))
    else:
        prof = list(pml.pml2(g, gl.fcut))
        if not top_damping:
            for q in (2, 3):
                prof[q][: nz // 2] = prof[q][nz // 2]
        prof = tuple(prof)
    o = Oracle2D(problem.Case2D(g, m, prof, gl, np.zeros((0, 2), np.int32), [], nt, 1, (0, 0, 0)), 64)
    off = nax + ext
    frames = []
    for it in range(1, nt + 1):
        o.step(it)
        if it % 5 == 0:
            frames.append([o.f[v][off + 3: off + 3 + nxu, 3: 3 + nzu].copy() for v in ("Vx", "Vz")])
    frames = np.array(frames)
    if ref is None:
        return frames
    return float(np.sqrt(((frames - ref) ** 2).sum() / (ref ** 2).sum()))


def test_pml_top_damping_makes_thicker_absorbers_reflect_more():
    """The reference's damping profile damps z within H of the top as well
    as the bottom (damping_profile, src/swpc_3d/m_absorb_p.f90:532-539, is
    called for zbeg too), and the side absorber columns start at k = 1
    (kbeg_a, m_global.f90:336-345) while the interior columns start at
    nz - na + 1.  Within na cells of the top the z damping therefore jumps
    from zero to d_z across each side interface, which reflects waves that
    travel along the top; the jump reaches deeper the thicker the
    absorber.  With that damping the error over the region grows with na
    (as the reference's absorber does); with the side columns undamped in
    z near the top it falls with na, as an absorber should."""
    ref = _region_error(0)
    with_top = {na: _region_error(na, True, ref=ref) for na in (6, 20, 60)}
    without = {na: _region_error(na, False, ref=ref) for na in (6, 20, 60)}
    print("reference profile", with_top, "no top z damping", without)
    assert with_top[6] < with_top[20] < with_top[60]
    assert without[6] > without[20] > without[60]
    assert without[60] < with_top[60] / 10
