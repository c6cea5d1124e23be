"""Case files for the Fortran hosts ``fhost/swpc3d_fdrive`` and
``fhost/swpc2d_fdrive``: the arrays a reference-style Fortran main program
holds after its setup phase, written in the layout the C ABI takes (see the
headers of fhost/swpc3d_fdrive.f90 and fhost/swpc2d_fdrive.f90)."""
from __future__ import annotations

import numpy as np

from .solver import PRECISION
from .source import ntw_of


def write_case3(path, case, precision="fp32"):
    g, m = case.grid, case.medium
    glut = case.glut
    st = np.asarray(case.stations, np.int32).reshape(-1, 3)
    nsrc = len(glut.idx)
    ntw = ntw_of(case.nt, case.ntdec_w)
    with open(path, "wb") as f:
        np.array([g.nx, g.ny, g.nz, g.na, case.nt, PRECISION[precision], nsrc, len(st), case.ntdec_w, ntw],
                 np.int32).tofile(f)
        np.array([g.dx, g.dy, g.dz, g.dt], np.float64).tofile(f)
        np.asarray(m.ts[:3], np.float32).tofile(f)
        for a in (m.bx, m.by, m.bz, m.lam, m.mu, m.taup, m.taus, m.muyz, m.muxz, m.muxy):
            np.ascontiguousarray(a, np.float32).tofile(f)
        for a in (m.kfs_top, m.kfs_bot, m.kob_top, m.kob_bot, m.kob):
            np.ascontiguousarray(a, np.int32).tofile(f)
        for a in case.pml:
            np.ascontiguousarray(a, np.float32).tofile(f)
        idx = np.asarray(glut.idx, np.int32).reshape(-1, 3)
        for q in range(3):
            np.ascontiguousarray(idx[:, q]).tofile(f)
        np.ascontiguousarray(glut.m, np.float64).tofile(f)
        np.ascontiguousarray(glut.table, np.float64).tofile(f)
        for q in range(3):
            np.ascontiguousarray(st[:, q]).tofile(f)
    return ntw, len(st)


def read_result3(path, ntw, nst):
    raw = open(path, "rb").read()
    n = 3 * nst * ntw
    vel = np.frombuffer(raw[:4 * n], np.float32).reshape(ntw, nst, 3)
    disp = np.frombuffer(raw[4 * n:12 * n], np.float64).reshape(ntw, nst, 3)
    vmax = np.frombuffer(raw[12 * n:12 * n + 12], np.float32)
    return vel, disp, vmax


def write_case2(path, case, precision="fp32", snaps=(), ntdec_s=1):
    """P-SV case; snaps holds (kind, ik (npts, 2), uc, m0) per stream."""
    g, m = case.grid, case.medium
    glut, bf = case.glut, case.bforce
    st = np.asarray(case.stations, np.int32).reshape(-1, 2)
    nsrc = len(glut.idx) if glut is not None else 0
    nbf = len(bf.idx) if bf is not None else 0
    ntw = ntw_of(case.nt, case.ntdec_w)
    with open(path, "wb") as f:
        np.array([g.nx, g.nz, g.na, case.nt, PRECISION[precision], nsrc, nbf, len(st), case.ntdec_w, ntw,
                  len(snaps), ntdec_s], np.int32).tofile(f)
        np.array([g.dx, g.dz, g.dt], np.float64).tofile(f)
        np.asarray(m.ts[:3], np.float32).tofile(f)
        for a in (m.rho, m.lam, m.mu, m.taup, m.taus):
            np.ascontiguousarray(a, np.float32).tofile(f)
        for a in (m.kfs_top, m.kfs_bot, m.kob_top, m.kob_bot, m.kob):
            np.ascontiguousarray(a, np.int32).tofile(f)
        for a in case.pml:
            np.ascontiguousarray(a, np.float32).tofile(f)
        for tab in (glut, bf):
            if tab is None or not len(tab.idx):
                continue
            idx = np.asarray(tab.idx, np.int32).reshape(-1, 2)
            for q in range(2):
                np.ascontiguousarray(idx[:, q]).tofile(f)
            np.ascontiguousarray(tab.m, np.float64).tofile(f)
            np.ascontiguousarray(tab.table, np.float64).tofile(f)
        for q in range(2):
            np.ascontiguousarray(st[:, q]).tofile(f)
        for kind, ik, uc, m0 in snaps:
            ik = np.ascontiguousarray(ik, np.int32).reshape(-1, 2)
            np.array([kind, len(ik)], np.int32).tofile(f)
            ik.tofile(f)
            np.array([uc, m0], np.float32).tofile(f)
    return ntw, len(st)


def read_result2(path, ntw, nst, snap_sizes=(), nframes=0):
    """vel (ntw, nst, 2), disp, vmax (2,), and per stream frames (nframes, 2, npts)."""
    raw = open(path, "rb").read()
    n = 2 * nst * ntw
    vel = np.frombuffer(raw[:4 * n], np.float32).reshape(ntw, nst, 2)
    disp = np.frombuffer(raw[4 * n:12 * n], np.float64).reshape(ntw, nst, 2)
    o = 12 * n
    vmax = np.frombuffer(raw[o:o + 8], np.float32)
    o += 8
    frames = []
    for npts in snap_sizes:
        k = 2 * npts * nframes
        frames.append(np.frombuffer(raw[o:o + 4 * k], np.float32).reshape(nframes, 2, npts))
        o += 4 * k
    assert o == len(raw), (o, len(raw))
    return vel, disp, vmax, frames
