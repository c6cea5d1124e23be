"""The SH oracle (osh_* in oracle/swpc_oracle.c, restating src/swpc_sh)
pinned by what the mathematics fixes: in a homogeneous elastic whole
space the 2-D SH line-force Green's function is

    G(r, t) = H(t - r/b) / (2 pi mu sqrt(t^2 - r^2/b^2)),

so a body force f s(t) gives u = f * int_0^acosh(bt/r) s(t - (r/b) cosh x) dx
/ (2 pi mu), and a moment source (myz, mxy) with moment function m(t) gives
u = -(mxy x/r + myz z/r) dG/dr * m.  The scheme's traces must follow
both (amplitude, arrival, radiation pattern and sign), which a dropped
term, a wrong sign, a transposed stencil or a wrong normalisation breaks.
"""
import math

import numpy as np
import pytest

from openswpc_amd import fdtool, medium as md, pml, problem
from openswpc_amd.problem import Case2D
from openswpc_amd.source import BodyForce, PointSource, Station, body_force_sh, station_cells2, stress_glut_sh
from oracle.oracle import OracleSH

NX = NZ = 300
DX, DT, NT = 0.05, 0.004, 650
TS, TR = 0.1, 0.5


def _whole_space():
    g = md.Grid2D(NX, NZ, DX, DX, DT, na=20, zbeg=2.0)    # zbeg > 0: no air, no free surface
    rho, lam, mu, qp, qs = md.layered2(g, [md.Layer(0, 2.7, 6.0, 3.5, 1e6, 1e6)])
    med = md.medium_setup2(g, rho, lam, mu, qp, qs, benchmark=True)   # elastic: tau = 0
    return g, med


def _stations(g, med, xs, zs):
    sts = [Station(xs + 2.5, 0.0, zs, "R1", "dep"), Station(xs + 1.5, 0.0, zs + 1.5, "R2", "dep"),
           Station(xs - 0.5, 0.0, zs - 2.0, "R3", "dep")]
    return station_cells2(g, med.kfs, med.kob, sts)


def _kupper(t):
    x = np.clip((t - TS) / TR, 0.0, 1.0)
    return np.where((t >= TS) & (t <= TS + TR), 3 * np.pi * np.sin(np.pi * x) ** 3 / (4 * TR), 0.0)


def _ikupper(t):
    c = np.cos(np.pi * np.clip((t - TS) / TR, 0.0, 1.0))
    return np.where(t < TS, 0.0, 0.75 * (2.0 / 3.0 - c + c ** 3 / 3.0))


def _line(r, t, b, mu, stf):
    """int_0^acosh(bt/r) stf(t - (r/b) cosh x) dx / (2 pi mu) for each t."""
    out = np.zeros_like(t)
    for q, tt in enumerate(t):
        if b * tt <= r:
            continue
        x = np.linspace(0.0, math.acosh(b * tt / r), 2001)
        out[q] = np.trapezoid(stf(tt - r / b * np.cosh(x)), x) / (2 * math.pi * mu)
    return out


def _geometry(g, med, idx, i, k):
    xc = lambda n: float(fdtool.i2x(n, g.xbeg, g.dx))   # noqa: E731
    zc = lambda n: float(fdtool.i2x(n, g.zbeg, g.dz))   # noqa: E731
    i0, k0 = idx
    mu = float(med.mu[i0 + 2, k0 + 2])
    b = math.sqrt(mu / float(med.rho[i0 + 2, k0 + 2]))
    return xc(i) - xc(i0), zc(k) - zc(k0), mu, b


def test_sh_body_force_matches_line_greens_function():
    g, med = _whole_space()
    xs, zs = 0.0, 2.0 + 150 * DX
    bf = body_force_sh(g, [BodyForce(xs, 0.0, zs, TS, TR, (0.0, 1.0, 0.0))], NT)
    cells, kept = _stations(g, med, xs, zs)
    o = OracleSH(Case2D(g, med, pml.pml_sh(g, bf.fcut), None, cells, kept, NT, 1, (0, 0, 0), bf), 64)
    o.run(1, NT)
    t = (np.arange(1, NT + 1) - 0.5) * DT
    for n, (i, k) in enumerate(cells):
        X, Z, mu, b = _geometry(g, med, bf.idx[0], i, k)
        u = float(bf.m[0]) * _line(math.hypot(X, Z), t, b, mu, _kupper)
        e = np.linalg.norm(o.disp[:, n] - u) / np.linalg.norm(u)
        print(f"station {n}: r = {math.hypot(X, Z):.3f} km, rel L2 {e:.2e}")
        assert e < 0.02


def test_sh_moment_source_matches_line_greens_function():
    g, med = _whole_space()
    xs, zs = 0.0, 2.0 + 150 * DX
    gl = stress_glut_sh(g, [PointSource(xs, 0.0, zs, TS, TR, 1e15, fdtool.sdr2moment(30.0, 60.0, 20.0))], NT)
    myz, mxy = gl.m[0]
    assert abs(myz) > 0.1 and abs(mxy) > 0.1
    cells, kept = _stations(g, med, xs, zs)
    o = OracleSH(Case2D(g, med, pml.pml_sh(g, gl.fcut), gl, cells, kept, NT, 1, (0, 0, 0), None), 64)
    o.run(1, NT)
    t = np.arange(1, NT + 1) * DT
    h = 1e-3
    for n, (i, k) in enumerate(cells):
        X, Z, mu, b = _geometry(g, med, gl.idx[0], i, k)
        r = math.hypot(X, Z)
        dG = (_line(r + h, t, b, mu, _ikupper) - _line(r - h, t, b, mu, _ikupper)) / (2 * h)
        u = -(mxy * X / r + myz * Z / r) * dG
        e = np.linalg.norm(o.disp[:, n] - u) / np.linalg.norm(u)
        print(f"station {n}: r = {r:.3f} km, rel L2 {e:.2e}")
        assert e < 0.01


def test_sh_pml_profile_uses_pa2():
    """The CFS frequency shift is a = a0 (1 - (x/H)^pa) with pa = 2 in the SH
    absorber and 1 in P-SV (m_absorb_p.f90 damping_profile).  a is
    recovered from the profile values g3 = (1 - h q)/(1 + h q), q = a + d/b,
    and g4 = (d/b)/(1 + h q), h = dt/2, at cells inside the absorber."""
    g = md.Grid2D(120, 80, 0.1, 0.1, 0.005, na=20)
    fcut = 1.0
    sh, psv = pml.pml_sh(g, fcut), pml.pml2(g, fcut)
    H = float(np.float32(g.na * g.dx))
    h = g.dt / 2
    for i in (3, 8, 14):
        x = float(g.xc[i + 2])
        r = ((g.xbeg + H) - x) / H
        for prof, pa in ((sh, 2), (psv, 1)):
            g1, g2, g3, g4 = (float(v) for v in prof[0][i - 1])
            q = (1 - g3) / (h * (1 + g3))
            a = q - g4 * (1 + h * q)
            assert a / (math.pi * fcut) == pytest.approx(1 - r ** pa, rel=2e-3, abs=2e-3), (i, pa)
    assert np.array_equal(sh[0][g.na + 5], psv[0][g.na + 5])      # interior: undamped either way


def test_sh_case_runs_finite():
    c = problem.case_sh(160, 100, 0.1, 0.005, 200, na=20, seed=3, rmed=(0.05, 2.0, 0.5), src_depth=4.0, nst=6,
                        zbeg=-1.0)
    o = OracleSH(c, 32)
    o.run(1, 200)
    assert np.isfinite(o.vel).all() and np.abs(o.vel).max() > 0
    assert o.vmax() > 0


def _sh_pml_case(nx, nz, na=5, dx=0.5, dt=0.04):
    g = md.Grid2D(nx, nz, dx, dx, dt, na=na, zbeg=-dx * 3)
    rho, lam, mu, qp, qs = md.layered2(g, [md.Layer(-1000.0, 2.7, 6.0, 3.5, 1e30, 1e30)])
    med = md.medium_setup2(g, rho, lam, mu, qp, qs, benchmark=True)
    return Case2D(g, med, pml.pml_sh(g, 1.0), stress_glut_sh(g, [], 2), np.zeros((0, 2), np.int32), [], 2, 1,
                  (0, 0, 0), None)


def _sh_box(g):
    return np.arange(-2, g.nx + 4)[:, None], np.arange(-2, g.nz + 4)[None, :]


def _sh_prof(g, prof, axis, q, at=None):
    """Profile column q over the (i, k) box along `axis`, or, with at =
    'i', the z profile indexed by i (broadcast along k)."""
    n = g.nz if (axis == 1 or at == "i") else g.nx
    v = np.full(n + 6, np.nan)
    v[3:3 + n] = prof[:, q - 1].astype(np.float64)
    if at == "i":
        w = np.full(g.nx + 6, np.nan)
        m = min(g.nx, g.nz)
        w[3:3 + m] = v[3:3 + m]
        return w.reshape(-1, 1)
    return v.reshape((n + 6, 1) if axis == 0 else (1, n + 6))


def _sh_mask(g):
    i, k = _sh_box(g)
    inside = (i >= 1) & (i <= g.nx) & (k >= 1) & (k <= g.nz)
    kbeg = np.where((i <= g.na) | (i >= g.nx - g.na + 1), 1, g.nz - g.na + 1)
    return inside & (k >= kbeg)


@pytest.mark.parametrize("nx, nz", [(24, 30), (30, 24)])
def test_sh_pml_velocity_closed_form(nx, nz):
    """Linear Sxy, Syz: the ADE values follow a_n = g4 c dt (1 + ... +
    g3^(n-1)) with gxc and gzc at k, and Vy advances by (g1 c + g2
    a_(n-1)) dt / rho per direction.  The z term of the velocity takes
    gzc(1:2, i) (src/swpc_sh/m_absorb_p.f90:180-181) where nx <= nz, the
    profile's bounds; with nx > nz that index runs past them and the
    profile at k is used."""
    case = _sh_pml_case(nx, nz)
    g = case.grid
    o = OracleSH(case, 64)
    o.med["rho"][...] = 2.5
    I, K = _sh_box(g)
    cxy, czy = 1.3, -0.8        # dSxy/dx, dSyz/dz (per km)
    o.set_field("Sxy", np.broadcast_to((cxy * I + 0.4 * K) * g.dx, g.shape))
    o.set_field("Syz", np.broadcast_to((-0.6 * I + czy * K) * g.dx, g.shape))
    gxc, gxe, gzc, gze = case.pml
    X = {q: _sh_prof(g, gxc, 0, q) for q in (1, 2, 3, 4)}
    Z = {q: _sh_prof(g, gzc, 1, q) for q in (1, 2, 3, 4)}
    Zv = {q: _sh_prof(g, gzc, 1, q, at="i") if nx <= nz else Z[q] for q in (1, 2)}
    dt = float(np.float32(g.dt))
    mask = _sh_mask(g)
    ax = az = 0.0
    v = np.zeros(g.shape)
    for step in range(1, 5):
        o.L.osh_absorb_vel(o._ps())
        v = v + float(np.float32(1) / np.float32(2.5)) * (X[1] * cxy + Zv[1] * czy + X[2] * ax + Zv[2] * az) * dt
        ax = X[4] * cxy * dt * sum(X[3] ** m for m in range(step))
        az = Z[4] * czy * dt * sum(Z[3] ** m for m in range(step))
        e = np.broadcast_to(v, g.shape)
        assert np.allclose(o.f["Vy"][mask], e[mask], rtol=1e-11, atol=1e-13), step
        assert np.all(o.f["Vy"][~mask] == 0)
        assert np.allclose(o.A["axSxy"][mask], np.broadcast_to(ax, g.shape)[mask], rtol=1e-11, atol=1e-15)
        assert np.allclose(o.A["azSyz"][mask], np.broadcast_to(az, g.shape)[mask], rtol=1e-11, atol=1e-15)


def test_sh_pml_stress_closed_form():
    """Linear Vy: the SH absorber advances its ADE values first and the
    stresses use the advanced ones (src/swpc_sh/m_absorb_p.f90:262-275):
    Sxy += mu_xy (gxe1 c + gxe2 a_n) dt with a_n = gxe4 c dt (1 + ... +
    gxe3^(n-1)), Syz likewise with gze; mu is uniform, so both harmonic
    means are 2 mu^2 / (2 mu + eps) in FP32."""
    case = _sh_pml_case(24, 30)
    g = case.grid
    o = OracleSH(case, 64)
    mu = 2.5
    o.med["mu"][...] = mu
    hm = float(np.float32(2 * mu * mu) / np.float32(2 * mu + np.float32(1.1920928955078125e-07)))
    I, K = _sh_box(g)
    cx, cz = 0.9, -1.7
    o.set_field("Vy", np.broadcast_to((cx * I + cz * K) * g.dx, g.shape))
    gxc, gxe, gzc, gze = case.pml
    X = {q: _sh_prof(g, gxe, 0, q) for q in (1, 2, 3, 4)}
    Z = {q: _sh_prof(g, gze, 1, q) for q in (1, 2, 3, 4)}
    dt = float(np.float32(g.dt))
    mask = _sh_mask(g)
    sxy = np.zeros(g.shape)
    syz = np.zeros(g.shape)
    for step in range(1, 5):
        o.L.osh_absorb_stress(o._ps())
        ax = X[4] * cx * dt * sum(X[3] ** m for m in range(step))
        az = Z[4] * cz * dt * sum(Z[3] ** m for m in range(step))
        sxy = sxy + hm * (X[1] * cx + X[2] * ax) * dt
        syz = syz + hm * (Z[1] * cz + Z[2] * az) * dt
        for name, e in (("Sxy", sxy), ("Syz", syz)):
            e = np.broadcast_to(e, g.shape)
            assert np.allclose(o.f[name][mask], e[mask], rtol=1e-6, atol=1e-13), (name, step)
            assert np.all(o.f[name][~mask] == 0)
