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


def _records(path):
    """Non-comment, non-blank lines split list-directed (blanks or commas)."""
    with open(path) as f:
        for line in f:
            t = line.strip()
            if not t or t.startswith("#"):
                continue
            yield t.replace(",", " ").split()


def _lonlat(lon, lat):
    if not (-360 <= lon <= 360 and -90 <= lat <= 90):
        raise ValueError(f"longitude/latitude ({lon}, {lat}) out of range")


def read_source_file(path, fmt="xym0ij", phi=0.0, clon=0.0, clat=0.0, mu_at=None, with_info=False):
    """Parse a moment-tensor source file (source__grid_moment,
    m_source.f90:467-699) in any of its formats: xym0ij, xym0dc, llm0ij,
    llm0dc, xymwij, xymwdc, llmwij, llmwdc, xydsdc, lldsdc, psmeca.

    Strike/dip/rake mechanisms use the strike measured from the grid
    azimuth phi; 'll' forms are mapped to the model frame about
    (clon, clat) with phi; the 'ds' forms take the moment as
    1e9 mu D S with mu the rigidity at the source cell, which ``mu_at(x, y, z)``
    supplies.  With with_info=True also returns the hypocentre record of
    the first source (evlo, evla, evdp, sx0, sy0, m6, otim)."""
    from . import geomap
    out = []
    info = None
    for w in _records(path):
        v = [float(x) for x in w[:12]] if fmt != "psmeca" else None
        if fmt in ("xym0ij", "xym0dc", "xymwij", "xymwdc", "xydsdc"):
            x, y = F32(v[0]), F32(v[1])
        elif fmt in ("llm0ij", "llm0dc", "llmwij", "llmwdc", "lldsdc"):
            _lonlat(v[0], v[1])
            x, y = geomap.g2c(v[0], v[1], clon, clat, phi)
        elif fmt == "psmeca":
            pass
        else:
            raise ValueError(f"unsupported stf_format {fmt!r}")
        if fmt == "psmeca":
            lon, lat, z = float(w[0]), float(w[1]), float(w[2])
            mzz, mxx, myy, mxz, myz, mxy = (F32(float(t)) for t in w[3:9])
            iex = int(w[9])
            myz, mxy = -myz, -mxy
            x, y = geomap.g2c(lon, lat, clon, clat, phi)
            # norm in dyn cm, then N m; rise time 2 x the empirical half duration
            s1 = mxx * mxx + myy * myy + mzz * mzz
            s2 = mxz * mxz + myz * myz + mxy * mxy
            m0tmp = float(np.sqrt(F32(s1 + F32(2) * s2)) / np.sqrt(F32(2.0)))
            mo = F32(m0tmp * float(F32(10.0 ** iex)))
            tr = F32(float(F32(F32(2) * F32(1.05)) * F32(1e-8)) * float(mo) ** (1.0 / 3.0))
            mo = F32(mo * F32(1e-7))
            m6 = tuple(F32(float(c) / m0tmp) for c in (mxx, myy, mzz, myz, mxz, mxy))
            s0 = PointSource(float(x), float(y), z, 0.0, float(tr), float(mo), m6)
        else:
            z, t0, tr, a = v[2], v[3], v[4], v[5]
            if fmt.endswith("ij"):
                m6 = tuple(F32(c) for c in v[6:12])
            else:
                m6 = _dc(*(v[7:10] if fmt.endswith("dsdc") else v[6:9]), phi)
            if fmt[2:4] == "m0":
                mo = F32(a)
            elif fmt[2:4] == "mw":
                mo = fdtool.seismic_moment(_mw(a))
            else:   # D [m] and S [km^2] times the cell rigidity
                if mu_at is None:
                    raise ValueError(f"stf_format {fmt} needs the medium (mu_at)")
                D, S = v[5], v[6]
                mo = F32(float(F32(F32(1e9) * F32(mu_at(float(x), float(y), z)))) * D * S)
            s0 = PointSource(float(x), float(y), z, t0, tr, float(mo), m6)
        out.append(s0)
        if info is None:
            from . import geomap as gm
            lo, la = gm.c2g(s0.x, s0.y, clon, clat, phi)
            info = {"evlo": float(lo), "evla": float(la), "evdp": float(F32(s0.z)), "sx0": s0.x, "sy0": s0.y,
                    "m6": tuple(float(c) for c in s0.m6), "otim": float(F32(s0.t0))}
    return (out, info) if with_info else out


def read_bodyforce_file(path, fmt="xym0ij", phi=0.0, clon=0.0, clat=0.0, with_info=False):
    """Body-force source file (source__grid_bodyforce, m_source.f90:701-773):
    x y z tstart trise fx fy fz, or lon lat in place of x y when the format
    starts with 'll'."""
    from . import geomap
    out = []
    info = None
    for w in _records(path):
        v = [float(t) for t in w[:8]]
        if fmt[:2] == "xy":
            x, y = F32(v[0]), F32(v[1])
        elif fmt[:2] == "ll":
            _lonlat(v[0], v[1])
            x, y = geomap.g2c(v[0], v[1], clon, clat, phi)
        else:
            raise ValueError(f"unsupported body-force stf_format {fmt!r}")
        b = BodyForce(float(x), float(y), v[2], v[3], v[4], (float(F32(v[5])), float(F32(v[6])), float(F32(v[7]))))
        out.append(b)
        if info is None:
            lo, la = geomap.c2g(b.x, b.y, clon, clat, phi)
            info = {"evlo": float(lo), "evla": float(la), "evdp": float(F32(b.z)), "sx0": b.x, "sy0": b.y,
                    "f3": b.f3, "otim": float(F32(b.t0))}
    return (out, info) if with_info else out


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


def stress_glut_sh(g, sources, nt, stftype="kupper"):
    """SH stress glut (src/swpc_sh/m_source.f90:554-594): m = (myz, mxy),
    the 4th and 6th moment-tensor components; table[it-1, n] =
    mo/M0 * momentrate(tbeg+(it-1/2)dt) * dt/(dx dz); the source y
    coordinate is ignored."""
    if not sources:
        return SourceTable(np.zeros((0, 2), np.int32), np.zeros((0, 2)), np.zeros((nt, 0)), 1.0, 0.0)
    t = stress_glut2(g, sources, nt, stftype)
    m2 = np.array([(s.m6[3], s.m6[5]) for s in sources], np.float64)
    return SourceTable(t.idx, m2, t.table, t.M0, t.fcut)


def body_force_sh(g, forces, nt, stftype="kupper"):
    """SH body force (src/swpc_sh/m_source.f90:596-629): fy normalised by
    M0 = sqrt(sum(fy)**2) = |sum fy| as the reference writes it (line 158);
    table[it-1, n] = momentrate(tbeg+it dt) * dt/(dx dz), no halving (one
    node); the kernel divides by rho(k, i)."""
    M0 = math.sqrt(float(F32(sum(float(F32(f.f3[1])) for f in forces))) ** 2)
    idx = [(fdtool.x2i(f.x, g.xbeg, g.dx), fdtool.x2i(f.z, g.zbeg, g.dz)) for f in forces]
    dt_dxz = float(F32(g.dt)) / (float(g.dx) * float(g.dz))
    tab = np.zeros((nt, len(forces)))
    for n, f in enumerate(forces):
        for it in range(1, nt + 1):
            tab[it - 1, n] = float(fdtool.momentrate(_t_vel(g.tbeg, it, g.dt), stftype, f.t0, f.tr)) * dt_dxz
    fy = np.array([float(F32(f.f3[1])) / M0 for f in forces])
    fcut = max(1.0 / float(F32(f.tr)) for f in forces)
    return SourceTable(np.array(idx, np.int32), fy, tab, M0, fcut)


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


def read_station_file(path, st_format="xy", clon=0.0, clat=0.0, phi=0.0, with_lonlat=False):
    """Station list (set_stinfo, m_wav.f90:144-190): 'xy' lines hold
    x y z name zsw, 'll' lines lon lat z name zsw; the other coordinate pair
    follows from the model frame about (clon, clat) rotated by phi.  With
    with_lonlat=True also returns the (lon, lat) of every station."""
    from . import geomap
    out, ll = [], []
    for w in _records(path):
        a, b, z = float(w[0]), float(w[1]), float(w[2])
        name, zsw = w[3][:8], (w[4] if len(w) > 4 else "dep")[:3]
        if st_format == "xy":
            x, y = float(F32(a)), float(F32(b))
            lo, la = geomap.c2g(x, y, clon, clat, phi)
        elif st_format == "ll":
            _lonlat(a, b)
            lo, la = F32(a), F32(b)
            x, y = (float(t) for t in geomap.g2c(a, b, clon, clat, phi))
        else:
            raise ValueError(f"unknown st_format: {st_format}")
        out.append(Station(x, y, float(F32(z)), name, zsw))
        ll.append((float(lo), float(la)))
    return (out, ll) if with_lonlat else out


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
