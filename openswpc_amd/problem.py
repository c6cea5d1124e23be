"""Synthetic problem builders for the benchmark configurations and tests.

A case bundles everything the hot path consumes: the grid, the assembled
medium, the PML profiles, the stress-glut source table and the receiver
cells.  Models follow the reference's recipes: layered media read like
``vmodel_lhm`` with an optional von Karman perturbation tiled like
``vmodel_lhm_rmed``; a double-couple Kupper source in the ``xym0dc`` form;
receivers on the free surface (``zsw = 'fsb'``).
"""
from __future__ import annotations

from dataclasses import dataclass
import math

import numpy as np

from . import fdtool, medium as md, pml, rmedia
from .source import (BodyForce, PointSource, Station, body_force2, body_force3, body_force_sh, station_cells2,
                     station_cells3, stress_glut2, stress_glut3, stress_glut_sh)

F32 = np.float32

# example/lhm.dat of the reference (depth rho vp vs qp qs)
LHM_EXAMPLE = [
    md.Layer(0, 2.3, 5.5, 3.14, 600, 300), md.Layer(3, 2.4, 6.0, 3.55, 600, 300),
    md.Layer(18, 2.8, 6.7, 3.83, 600, 300), md.Layer(33, 3.2, 7.8, 4.46, 600, 300),
    md.Layer(100, 3.3, 8.0, 4.57, 600, 300), md.Layer(225, 3.4, 8.4, 4.80, 600, 300),
    md.Layer(325, 3.5, 8.6, 4.91, 600, 300), md.Layer(425, 3.7, 9.3, 5.31, 600, 300),
]

# two flat layers of the P-SV configuration
LHM_TWO = [md.Layer(0, 2.7, 5.8, 3.36, 600, 300), md.Layer(15, 2.9, 6.5, 3.75, 600, 300)]


@dataclass
class Case3D:
    grid: md.Grid3D
    medium: md.Medium3D
    pml: tuple
    glut: object
    stations: np.ndarray
    station_list: list
    nt: int
    ntdec_w: int
    mech: tuple
    bforce: object = None     # body-force table (bf_mode) instead of the stress glut


@dataclass
class Case2D:
    grid: md.Grid2D
    medium: md.Medium2D
    pml: tuple
    glut: object
    stations: np.ndarray
    station_list: list
    nt: int
    ntdec_w: int
    mech: tuple
    bforce: object = None


def mechanism(seed):
    """Strike, dip and rake drawn from the seed: U[0,360), U[0,90], U[-180,180)."""
    rng = np.random.default_rng(seed)
    return float(rng.uniform(0, 360)), float(rng.uniform(0, 90)), float(rng.uniform(-180, 180))


def case3d(nx, ny, nz, dx, dt, nt, na=20, layers=LHM_EXAMPLE, seed=1, rmed=None, src_depth=10.0,
           tr=1.0, t0=0.1, mo=1e15, nst_side=8, st_spacing=None, ntdec_w=1, rmed_tile=(128, 128, 64),
           benchmark=False, zbeg=None, bodyforce=False, src_xy=(0.0, 0.0), log=None, device_medium=False):
    """3-D case: layered (+ random) medium, double couple at src_depth below
    the grid centre, an nst_side x nst_side receiver grid on the free surface.
    bodyforce=True replaces the double couple by a single force of the same
    position and time function with a direction drawn from the seed (the
    reference's bf_mode).

    rmed = (epsilon, a_km, kappa) adds a von Karman perturbation generated
    on a periodic tile and mapped by rmed_tile periodicity.  log (optional)
    receives a line after each setup stage.

    device_medium=True leaves the random-medium synthesis and medium__setup
    to the GPU: ``medium`` is then a RawMedium3D, and the receivers (``stations`` None, ``station_list``
    the Station list) are placed by ``open_case3d`` from the kfs/kob maps
    the device setup returns.
    """
    import time
    clock0 = time.time()
    note = (lambda m: log(f"case3d: {m} ({time.time() - clock0:.0f} s)")) if log else (lambda m: None)
    g = md.Grid3D(nx, ny, nz, dx, dx, dx, dt, na=na, zbeg=zbeg)
    xi = None
    if rmed is not None:
        eps, a, kappa = rmed
        tile = tuple(min(t, n) for t, n in zip(rmed_tile, (nx, ny, nz)))
        vol = rmedia.generate(tile, dx, a, eps, kappa, seed=seed, device=device_medium)
        xi = md.rmed_tile(vol, -2, nx + 3, -2, ny + 3, -2, nz + 3)
        note("random medium")
    rho, lam, mu, qp, qs = md.layered3(g, layers, xi=xi)
    xi = None
    note("layered model")
    if device_medium:
        med = md.RawMedium3D(rho, lam, mu, qp, qs, benchmark=benchmark)
    else:
        med = md.medium_setup3(g, rho, lam, mu, qp, qs, benchmark=benchmark)
        note("medium__setup")
    strike, dip, rake = mechanism(seed)
    src = PointSource(float(src_xy[0]), float(src_xy[1]), src_depth, t0, tr, mo, fdtool.sdr2moment(strike, dip, rake))
    glut = stress_glut3(g, [src], nt)
    bforce = None
    if bodyforce:
        f = np.random.default_rng(seed + 1000).standard_normal(3)
        bforce = body_force3(g, [BodyForce(0.0, 0.0, src_depth, t0, tr, tuple(float(v) for v in f))], nt)
        glut = None
    prof = pml.pml3(g, (bforce or glut).fcut)
    if st_spacing is None:
        st_spacing = (nx - 2 * na) * dx / (nst_side + 1)
    sts = []
    for a_ in range(nst_side):
        for b_ in range(nst_side):
            x = (a_ - (nst_side - 1) / 2) * st_spacing
            y = (b_ - (nst_side - 1) / 2) * st_spacing
            sts.append(Station(x + 0.3 * dx, y + 0.3 * dx, 0.0, f"S{a_:02d}{b_:02d}", "fsb"))
    if device_medium:
        return Case3D(g, med, prof, glut, None, sts, nt, ntdec_w, (strike, dip, rake), bforce)
    cells, kept = station_cells3(g, med.kfs, med.kob, sts)
    return Case3D(g, med, prof, glut, cells, kept, nt, ntdec_w, (strike, dip, rake), bforce)


def open_case3d(case, precision="fp32"):
    """A Swpc3D for a case; with a RawMedium3D (device_medium) the medium is
    set up on the GPU and the receivers are placed from its kfs/kob maps
    (case.stations and case.station_list are updated)."""
    from .solver import Swpc3D
    raw = isinstance(case.medium, md.RawMedium3D)
    sim = Swpc3D(case.grid, case.medium, case.pml, case.glut, None if raw else case.stations, case.ntdec_w, case.nt,
                 precision=precision, bodyforce=case.bforce)
    if raw:
        if case.stations is None:
            case.stations, case.station_list = station_cells3(case.grid, case.medium.kfs, case.medium.kob,
                                                              case.station_list)
        if len(case.stations):
            sim.set_receivers(case.stations)
    return sim


def case2d(nx, nz, dx, dt, nt, na=20, layers=LHM_TWO, seed=1, rmed=None, src_depth=10.0, tr=1.0, t0=0.1,
           mo=1e15, nst=16, st_spacing=2.0, ntdec_w=1, benchmark=False, zbeg=None, bodyforce=False, src_x=0.0):
    """P-SV case: layered (+ random) medium, double couple (its x-z part) at
    src_depth, nst receivers on the free surface at st_spacing km."""
    g = md.Grid2D(nx, nz, dx, dx, dt, na=na, zbeg=zbeg)
    xi = None
    if rmed is not None:
        eps, a, kappa = rmed
        vol = rmedia.generate((nx + 6, nz + 6), dx, a, eps, kappa, seed=seed)
        xi = vol
    rho, lam, mu, qp, qs = md.layered2(g, layers, xi=xi)
    med = md.medium_setup2(g, rho, lam, mu, qp, qs, benchmark=benchmark)
    strike, dip, rake = mechanism(seed)
    src = PointSource(float(src_x), 0.0, src_depth, t0, tr, mo, fdtool.sdr2moment(strike, dip, rake))
    glut = stress_glut2(g, [src], nt)
    bforce = None
    if bodyforce:
        f = np.random.default_rng(seed + 1000).standard_normal(3)
        bforce = body_force2(g, [BodyForce(0.0, 0.0, src_depth, t0, tr, tuple(float(v) for v in f))], nt)
        glut = None
    prof = pml.pml2(g, (bforce or glut).fcut)
    sts = [Station((n - (nst - 1) / 2) * st_spacing + 0.3 * dx, 0.0, 0.0, f"P{n:02d}", "fsb") for n in range(nst)]
    cells, kept = station_cells2(g, med.kfs, med.kob, sts)
    return Case2D(g, med, prof, glut, cells, kept, nt, ntdec_w, (strike, dip, rake), bforce)


def case_sh(nx, nz, dx, dt, nt, na=20, layers=LHM_TWO, seed=1, rmed=None, src_depth=10.0, tr=1.0, t0=0.1,
            mo=1e15, nst=16, st_spacing=2.0, ntdec_w=1, benchmark=False, zbeg=None, bodyforce=False, src_x=0.0):
    """SH case on the x-z plane (swpc_sh): the P-SV medium recipe (the SH
    medium__setup is the P-SV one, src/swpc_sh/m_medium.f90), a double
    couple whose (myz, mxy) components drive the SH stress glut, nst
    receivers on the free surface; bodyforce=True uses a single force fy
    of the same time function instead."""
    g = md.Grid2D(nx, nz, dx, dx, dt, na=na, zbeg=zbeg)
    xi = None
    if rmed is not None:
        eps, a, kappa = rmed
        xi = rmedia.generate((nx + 6, nz + 6), dx, a, eps, kappa, seed=seed)
    rho, lam, mu, qp, qs = md.layered2(g, layers, xi=xi)
    med = md.medium_setup2(g, rho, lam, mu, qp, qs, benchmark=benchmark)
    strike, dip, rake = mechanism(seed)
    src = PointSource(float(src_x), 0.0, src_depth, t0, tr, mo, fdtool.sdr2moment(strike, dip, rake))
    glut = stress_glut_sh(g, [src], nt)
    bforce = None
    if bodyforce:
        fy = float(np.random.default_rng(seed + 1000).standard_normal())
        bforce = body_force_sh(g, [BodyForce(float(src_x), 0.0, src_depth, t0, tr, (0.0, fy, 0.0))], nt)
        glut = None
    prof = pml.pml_sh(g, (bforce or glut).fcut)
    sts = [Station((n - (nst - 1) / 2) * st_spacing + 0.3 * dx, 0.0, 0.0, f"H{n:02d}", "fsb") for n in range(nst)]
    cells, kept = station_cells2(g, med.kfs, med.kob, sts)
    return Case2D(g, med, prof, glut, cells, kept, nt, ntdec_w, (strike, dip, rake), bforce)


def stability3(g, vmax):
    """Courant number of the 4th-order scheme as in report__setup
    (src/swpc_3d/m_report.f90:71-104): c = vmax dt / dh * 7/6."""
    dh = 1.0 / math.sqrt(1 / g.dx ** 2 + 1 / g.dy ** 2 + 1 / g.dz ** 2)
    return vmax * g.dt / dh * 7.0 / 6.0
