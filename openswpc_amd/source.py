"""Sources and stations: file formats, grid mapping and the per-step source
tables handed to the kernels.

Restates the host side of ``source__setup`` / ``source__grid_moment``
(src/swpc_3d/m_source.f90:41-313, 467-560; P-SV src/swpc_psv/m_source.f90)
and ``set_stinfo`` (src/swpc_3d/m_wav.f90:144-274).
"""
from __future__ import annotations

from dataclasses import dataclass
import math

import numpy as np

from . import fdtool

F32 = np.float32


@dataclass
class PointSource:
    x: float
    y: float
    z: float
    t0: float          # rupture start time srcprm(1)
    tr: float          # rise time srcprm(2)
    mo: float          # seismic moment [N m]
    m6: tuple          # (mxx, myy, mzz, myz, mxz, mxy)


@dataclass
class BodyForce:
    x: float
    y: float
    z: float
    t0: float
    tr: float
    f3: tuple          # (fx, fy, fz)


def read_source_file(path, fmt="xym0ij", phi=0.0):
    """Parse a source file in the xym0ij / xym0dc / xymwij / xymwdc formats
    (m_source.f90:511-560).  Mechanisms given as strike/dip/rake use the
    strike measured from the grid azimuth phi."""
    out = []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            w = [float(v) for v in s.replace(",", " ").split()]
            x, y, z, t0, tr = w[:5]
            if fmt == "xym0ij":
                mo, m6 = w[5], tuple(F32(v) for v in w[6:12])
            elif fmt == "xym0dc":
                mo = w[5]
                m6 = _dc(*w[6:9], phi)
            elif fmt == "xymwij":
                mo, m6 = float(fdtool.seismic_moment(_mw(w[5]))), tuple(F32(v) for v in w[6:12])
            elif fmt == "xymwdc":
                mo = float(fdtool.seismic_moment(_mw(w[5])))
                m6 = _dc(*w[6:9], phi)
            else:
                raise ValueError(f"unsupported stf_format {fmt!r}")
            out.append(PointSource(x, y, z, t0, tr, float(F32(mo)), m6))
    return out


def _mw(mw):
    if mw > 11.0:
        raise ValueError("moment magnitude must be <= 11")
    return mw


def _dc(strike, dip, rake, phi):
    if not (-360 <= strike <= 360 and -90 <= dip <= 90 and -180 <= rake <= 180):
        raise ValueError("strike/dip/rake out of range")
    return fdtool.sdr2moment(F32(strike) - F32(phi), dip, rake)


def write_source_file(path, sources, fmt="xym0ij"):
    with open(path, "w") as f:
        f.write("# x y z tstart trise mo mxx myy mzz myz mxz mxy\n")
        for s in sources:
            f.write(" ".join(f"{v:.8g}" for v in (s.x, s.y, s.z, s.t0, s.tr, s.mo, *s.m6)) + "\n")


@dataclass
class SourceTable:
    idx: np.ndarray      # (nsrc, 3) or (nsrc, 2): 1-based cell indices
    m: np.ndarray        # (nsrc, 6) or (nsrc, 3): moment components
    table: np.ndarray    # (nt, nsrc): per-step scalar handed to the kernels
    M0: float            # total moment, multiplied back at output
    fcut: float          # 1 / min rise time (m_source.f90:203-207)


def _t_stress(tbeg, it, dt):
    # tbeg + (it - 0.5) * dt evaluated in real(SP)
    return F32(F32(tbeg) + (F32(it) - F32(0.5)) * F32(dt))


def _t_vel(tbeg, it, dt):
    return F32(F32(tbeg) + F32(it) * F32(dt))


def stress_glut3(g, sources, nt, stftype="kupper"):
    """Source table for source__stressglut: table[it-1, n] =
    (mo/M0) * momentrate(tbeg+(it-1/2)dt) * dt/(dx dy dz), in FP64."""
    if not sources:
        return SourceTable(np.zeros((0, 3), np.int32), np.zeros((0, 6)), np.zeros((nt, 0)), 1.0, 0.0)
    M0 = float(sum(float(s.mo) for s in sources))
    idx = []
    for s in sources:
        i = fdtool.x2i(s.x, g.xbeg, g.dx)
        j = fdtool.x2i(s.y, g.ybeg, g.dy)
        k = fdtool.x2i(s.z, g.zbeg, g.dz)
        if not (g.xbeg <= s.x <= g.xend and g.ybeg <= s.y <= g.yend and g.zbeg <= s.z <= g.zend):
            raise ValueError(f"source at ({s.x}, {s.y}, {s.z}) outside the model (m_source.f90:276-280)")
        idx.append((i, j, k))
    dt_dxyz = float(F32(g.dt)) / (float(g.dx) * float(g.dy) * float(g.dz))
    tab = np.zeros((nt, len(sources)), np.float64)
    for n, s in enumerate(sources):
        mon = float(s.mo) / M0
        for it in range(1, nt + 1):
            st = fdtool.momentrate(_t_stress(g.tbeg, it, g.dt), stftype, s.t0, s.tr)
            tab[it - 1, n] = mon * float(st) * dt_dxyz
    fcut = max(1.0 / float(F32(s.tr)) for s in sources)
    return SourceTable(np.array(idx, np.int32), np.array([s.m6 for s in sources], np.float64), tab, M0, fcut)


def stress_glut2(g, sources, nt, stftype="kupper"):
    """P-SV stress glut (src/swpc_psv/m_source.f90:572-614): m = (mxx, mzz, mxz),
    table uses dt/(dx dz); the source y coordinate is ignored."""
    if not sources:
        return SourceTable(np.zeros((0, 2), np.int32), np.zeros((0, 3)), np.zeros((nt, 0)), 1.0, 0.0)
    M0 = float(sum(float(s.mo) for s in sources))
    idx = []
    for s in sources:
        i = fdtool.x2i(s.x, g.xbeg, g.dx)
        k = fdtool.x2i(s.z, g.zbeg, g.dz)
        if not (g.xbeg <= s.x <= g.xend and g.zbeg <= s.z <= g.zend):
            raise ValueError(f"source at ({s.x}, {s.z}) outside the model")
        idx.append((i, k))
    dt_dxz = float(F32(g.dt)) / (float(g.dx) * float(g.dz))
    tab = np.zeros((nt, len(sources)), np.float64)
    for n, s in enumerate(sources):
        mon = float(s.mo) / M0
        for it in range(1, nt + 1):
            st = fdtool.momentrate(_t_stress(g.tbeg, it, g.dt), stftype, s.t0, s.tr)
            tab[it - 1, n] = mon * float(st) * dt_dxz
    m3 = np.array([(s.m6[0], s.m6[2], s.m6[4]) for s in sources], np.float64)
    fcut = max(1.0 / float(F32(s.tr)) for s in sources)
    return SourceTable(np.array(idx, np.int32), m3, tab, M0, fcut)


def body_force3(g, forces, nt, stftype="kupper"):
    """Body-force table (m_source.f90:835-876): f3 normalised by
    M0 = |f| (m_source.f90:212-216); table[it-1, n] = momentrate(tbeg+it dt) dt/(dx dy dz)/2."""
    M0 = math.sqrt(sum(float(F32(f.f3[0])) ** 2 + float(F32(f.f3[1])) ** 2 + float(F32(f.f3[2])) ** 2 for f in forces))
    idx = [(fdtool.x2i(f.x, g.xbeg, g.dx), fdtool.x2i(f.y, g.ybeg, g.dy), fdtool.x2i(f.z, g.zbeg, g.dz)) for f in forces]
    dt_dxyz = float(F32(g.dt)) / (float(g.dx) * float(g.dy) * float(g.dz))
    tab = np.zeros((nt, len(forces)))
    for n, f in enumerate(forces):
        for it in range(1, nt + 1):
            tab[it - 1, n] = float(fdtool.momentrate(_t_vel(g.tbeg, it, g.dt), stftype, f.t0, f.tr)) * dt_dxyz / 2
    f3 = np.array([[float(v) / M0 for v in f.f3] for f in forces])
    fcut = max(1.0 / float(F32(f.tr)) for f in forces)
    return SourceTable(np.array(idx, np.int32), f3, tab, M0, fcut)


def body_force2(g, forces, nt, stftype="kupper"):
    """P-SV body-force table (src/swpc_psv/m_source.f90:616-654): f2 =
    (fx, fz) normalised by M0 = |f| (m_source.f90:157-158); table[it-1, n] =
    momentrate(tbeg+it dt) dt/(dx dz)/2; the kernel divides by rho(k,i)."""
    M0 = math.sqrt(sum(float(F32(f.f3[0])) ** 2 + float(F32(f.f3[2])) ** 2 for f in forces))
    idx = [(fdtool.x2i(f.x, g.xbeg, g.dx), fdtool.x2i(f.z, g.zbeg, g.dz)) for f in forces]
    dt_dxz = float(F32(g.dt)) / (float(g.dx) * float(g.dz))
    tab = np.zeros((nt, len(forces)))
    for n, f in enumerate(forces):
        for it in range(1, nt + 1):
            tab[it - 1, n] = float(fdtool.momentrate(_t_vel(g.tbeg, it, g.dt), stftype, f.t0, f.tr)) * dt_dxz / 2
    f2 = np.array([[float(f.f3[0]) / M0, float(f.f3[2]) / M0] for f in forces])
    fcut = max(1.0 / float(F32(f.tr)) for f in forces)
    return SourceTable(np.array(idx, np.int32), f2, tab, M0, fcut)


# ---------------------------------------------------------------------------
# stations
# ---------------------------------------------------------------------------
@dataclass
class Station:
    x: float
    y: float
    z: float
    name: str
    zsw: str = "dep"


def read_station_file(path):
    """Station list in the 'xy' format: x y z name zsw (m_wav.f90:176-180)."""
    out = []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            w = s.split()
            out.append(Station(float(w[0]), float(w[1]), float(w[2]), w[3], w[4] if len(w) > 4 else "dep"))
    return out


def ntw_of(nt, ntdec_w):
    """ntw = floor(real(nt-1)/real(ntdec_w) + 1) (m_wav.f90:92)."""
    return int(math.floor(float(F32(F32(nt - 1) / F32(ntdec_w) + F32(1.0)))))


def station_cells3(g, kfs, kob, stations):
    """(ist, jst, kst, kept) per set_stinfo (m_wav.f90:196-248) for one domain."""
    out, kept = [], []
    x1, xn = fdtool.i2x(1, g.xbeg, g.dx), fdtool.i2x(g.nx, g.xbeg, g.dx)
    y1, yn = fdtool.i2x(1, g.ybeg, g.dy), fdtool.i2x(g.ny, g.ybeg, g.dy)
    for s in stations:
        i = fdtool.x2i(s.x, g.xbeg, g.dx)
        j = fdtool.x2i(s.y, g.ybeg, g.dy)
        k = fdtool.x2i(s.z, g.zbeg, g.dz)
        if not (x1 < F32(s.x) < xn and y1 < F32(s.y) < yn and 1 < k < g.nz):
            continue
        if s.zsw == "fsb":
            k = int(kfs[j + 2, i + 2]) + 1
        elif s.zsw == "obb":
            k = int(kob[j + 2, i + 2]) + 1
        elif s.zsw == "oba":
            k = int(kob[j + 2, i + 2]) - 1
        if k > g.nz:
            k = g.nz - 1
        if k < 1:
            k = 2
        out.append((i, j, k))
        kept.append(s)
    return np.array(out, np.int32).reshape(-1, 3), kept


def station_cells2(g, kfs, kob, stations):
    """P-SV twin of station_cells3: stations on the x-z plane."""
    out, kept = [], []
    x1, xn = fdtool.i2x(1, g.xbeg, g.dx), fdtool.i2x(g.nx, g.xbeg, g.dx)
    for s in stations:
        i = fdtool.x2i(s.x, g.xbeg, g.dx)
        k = fdtool.x2i(s.z, g.zbeg, g.dz)
        if not (x1 < F32(s.x) < xn and 1 < k < g.nz):
            continue
        if s.zsw == "fsb":
            k = int(kfs[i + 2]) + 1
        elif s.zsw == "obb":
            k = int(kob[i + 2]) + 1
        elif s.zsw == "oba":
            k = int(kob[i + 2]) - 1
        if k > g.nz:
            k = g.nz - 1
        if k < 1:
            k = 2
        out.append((i, k))
        kept.append(s)
    return np.array(out, np.int32).reshape(-1, 2), kept
