"""Run a simulation described by a reference-style ``input.inf`` on the GPU.

This is the host program of the reference (src/swpc_3d/main.f90:36-160;
P-SV src/swpc_psv/main.f90) around the GPU time loop:

* parameters: global__readprm (src/swpc_3d/m_global.f90:125-178), with
  benchmark_mode's fixed values;
* medium: medium__setup (m_medium.f90:36-130) with the velocity models
  'uni', 'uni_rmed', 'lhm' and 'lhm_rmed' (m_vmodel_uni.f90,
  m_vmodel_uni_rmed.f90, m_vmodel_lhm.f90, m_vmodel_lhm_rmed.f90), random
  media read from classic netCDF volumes (src/shared/m_rdrmed.f90) and
  the Munk sea-water profile;
* sources: source__setup (m_source.f90:41-313), every stf_format of
  source__grid_moment / source__grid_bodyforce, bf_mode, sdep_fit 'bd0',
  and the benchmark source;
* stations: wav__setup / set_stinfo (m_wav.f90:54-274), 'xy' and 'll';
* snapshots: snap__setup (m_snap.f90:97-323), native or netCDF files;
* report__setup's stability check (m_report.f90:40-118) and
  report__progress (m_report.f90:120-185) with its divergence check;
* outputs: wav__write in the 'sac', 'csf', 'tar_st' and 'tar_node'
  formats, the progressive write of ntdec_w_prg, snap__closefiles.

Options outside this host raise InputError: the velocity models read from
grid files ('grd', 'grd_rmed', 'lgm', 'lgm_rmed', 'user'), pw_mode,
green_mode, stabilize_pml, earth_flattening, sw_wav_stress,
sw_wav_strain, and absorbers other than 'pml'.  nproc_x/nproc_y and the
pads are read and ignored: one GPU holds the whole model.

``python -m openswpc_amd.driver -i input.inf [-p fp32|fp64|fp32_r16|fp32_s16]``
"""
from __future__ import annotations

import argparse
import math
import os
import sys
import time
from dataclasses import dataclass, field

import numpy as np

from . import cdf, fdtool, medium as md, pml, problem, sac, snapshot
from .inifile import InputFile
from .source import (BodyForce, PointSource, body_force2, body_force3, body_force_sh, read_bodyforce_file,
                     read_source_file, read_station_file, station_cells2, station_cells3, stress_glut2, stress_glut3,
                     stress_glut_sh)

F32 = np.float32
SNAP3 = [(p, k) for k in ("ps", "v", "u") for p in ("yz", "xz", "xy", "fs", "ob")]


class InputError(ValueError):
    """An input.inf setting this host does not run, or an invalid one."""


def _ini(inp):
    if isinstance(inp, InputFile):
        return inp
    return InputFile(inp)


@dataclass
class Setup:
    """Everything the time loop and the outputs need."""
    dim: str
    ini: InputFile
    title: str
    odir: str
    nt: int
    grid: object
    medium: object
    prof: tuple
    glut: object
    bforce: object
    stations: list
    cells: np.ndarray
    stlonlat: list
    ntdec_w: int
    ntdec_r: int
    wav: dict
    snaps: list
    snap_cfg: dict
    meta: dict
    clon: float
    clat: float
    phi: float
    bddep0: object
    report: dict = field(default_factory=dict)


# ---------------------------------------------------------------------------
# parameters
# ---------------------------------------------------------------------------
def _unsupported(ini, dim):
    checks = [("pw_mode", False, "plane-wave mode (pw_mode)"),
              ("stabilize_pml", False, "stabilize_pml"), ("earth_flattening", False, "earth_flattening"),
              ("sw_wav_stress", False, "stress waveforms (sw_wav_stress)"),
              ("sw_wav_strain", False, "strain waveforms (sw_wav_strain)")]
    if ini.get("green_mode", False):
        # wav__setup empties the station list in green_mode (m_wav.f90:66-73)
        # before green__setup looks the pseudo-source station up in it
        # (m_green.f90:159, m_wav.f90:643), so the reference stops at
        # assert(src_somewhere) (m_green.f90:163) for every input
        raise InputError("Green's function mode (green_mode): the reference stops in green__setup for every input "
                         "(the station list is empty when the pseudo-source is looked up); not supported")
    for key, dflt, what in checks:
        if ini.get(key, dflt):
            raise InputError(f"{what} is not supported by this host")
    if ini.get("abc_type", "pml") != "pml":
        raise InputError("only abc_type = 'pml' is supported")


def _grid(ini, dim):
    bench = ini.get("benchmark_mode", False)
    nx = ini.get("nx", 256)
    ny = ini.get("ny", 256) if dim == "3d" else 1
    nz = ini.get("nz", 256)
    if bench:
        dx = dy = dz = 0.5
        dt = F32(0.04)
        na = 20
        xbeg = float(F32(-nx / 2.0 * float(F32(dx))))
        ybeg = float(F32(-ny / 2.0 * float(F32(dy))))
        zbeg = float(F32(-30 * F32(dz)))
        tbeg, clon, clat, phi = 0.0, F32(139.7604), F32(35.7182), F32(0.0)
    else:
        dx = ini.get("dx", 0.5)
        dy = ini.get("dy", 0.5) if dim == "3d" else dx
        dz = ini.get("dz", 0.5)
        dt = ini.get("dt", 0.01, "s")
        na = ini.get("na", 20)
        xbeg = float(ini.get("xbeg", F32(-(nx // 2) * F32(dx)), "s"))
        ybeg = float(ini.get("ybeg", F32(-(ny // 2) * F32(dy)), "s")) if dim == "3d" else None
        zbeg = float(ini.get("zbeg", F32(-30 * F32(dz)), "s"))
        tbeg = float(ini.get("tbeg", 0.0, "s"))
        clon = ini.get("clon", 139.7604, "s")
        clat = ini.get("clat", 35.7182, "s")
        phi = ini.get("phi", 0.0, "s")
    if dim == "3d":
        g = md.Grid3D(nx, ny, nz, dx, dy, dz, float(dt), na=na, xbeg=xbeg, ybeg=ybeg, zbeg=zbeg, tbeg=tbeg)
    else:
        g = md.Grid2D(nx, nz, dx, dz, float(dt), na=na, xbeg=xbeg, zbeg=zbeg, tbeg=tbeg)
    return g, bench, float(clon), float(clat), float(phi)


# ---------------------------------------------------------------------------
# velocity models
# ---------------------------------------------------------------------------
def read_rmed(path, ndim):
    """A random-medium volume of gen_rmed3d/gen_rmed2d: the data variable
    (the 4th variable in 3-D, the 3rd in 2-D, as rdrmed__3d/2d pick it) as
    (nxc, nyc, nzc) or (nxc, nzc) with x first."""
    ds = cdf.read(path)
    names = list(ds.variables)
    v = ds.variables[names[3 if ndim == 3 else 2]]
    return np.ascontiguousarray(v.data().T)


def _rmed_field(g, dim, path):
    if not os.path.exists(path):
        return None
    vol = read_rmed(path, 3 if dim == "3d" else 2)
    if dim == "3d":
        return md.rmed_tile(vol, -2, g.nx + 3, -2, g.ny + 3, -2, g.nz + 3)
    return md.rmed_tile2(vol, -2, g.nx + 3, -2, g.nz + 3)


def velocity_model(ini, g, dim, bench):
    """(rho, lam, mu, qp, qs, bddep0): medium over the memory box and the
    depth of boundary 0 (the topography) of the vmodel_* routines."""
    shape = g.shape
    zc = g.zc
    if bench:
        # benchmark medium (m_medium.f90:52-74): air above z = 0, a Poisson solid below
        z = np.broadcast_to(np.asarray(zc, F32), shape)
        air = z < 0
        rho = np.where(air, F32(0.001), F32(2.7)).astype(F32)
        mu = np.where(air, F32(0), F32(F32(2.7) * F32(3.5) * F32(3.5))).astype(F32)
        lam = mu.copy()
        big = np.full(shape, F32(1e10))
        return rho, lam, mu, big, big.copy(), 0.0
    vtype = ini.get("vmodel_type", "uni")
    vcut = float(ini.get("vcut", 0.0, "s"))
    munk = ini.get("munk_profile", False)
    vmax = md.vmax_clip3(g) if dim == "3d" else md.vmax_clip2(g)
    if vtype in ("uni", "uni_rmed"):
        vp0 = ini.get("vp0", 5.0, "s")
        vs0 = ini.get("vs0", F32(vp0 / F32(math.sqrt(3.0))), "s")
        rho0 = ini.get("rho0", 2.7, "s")
        qp0 = ini.get("qp0", 1000000.0, "s")
        qs0 = ini.get("qs0", 1000000.0, "s")
        topo0 = ini.get("topo0", 0.0, "s")
        xi = None
        rhomin = 1.0
        if vtype == "uni_rmed":
            rhomin = float(ini.get("rhomin", 1.0, "s"))
            path = os.path.join(ini.get("dir_rmed", ""), ini.get("fn_rmed0", ""))
            xi = _rmed_field(g, dim, path)
            if xi is None:
                print(f"[info] rmedia file {path} not found", file=sys.stderr)
                xi = np.zeros(shape, F32)
        out = md.vmodel_uni(zc, shape, vp0, vs0, rho0, qp0, qs0, topo0, xi, vcut, vmax, rhomin, munk)
        return (*out, float(topo0))
    if vtype in ("lhm", "lhm_rmed"):
        if vtype == "lhm":
            fn = ini.get("fn_lhm", "")
            if not os.path.exists(fn):
                raise InputError(f"velocity model file {fn!r} not found")
            layers = md.read_lhm(fn)
            cols = md.vmodel_lhm(zc, layers, vcut, munk=munk)
            out = tuple(np.ascontiguousarray(np.broadcast_to(c, shape)).astype(F32) for c in cols)
        else:
            fn = ini.get("fn_lhm_rmed", "")
            if not os.path.exists(fn):
                raise InputError(f"velocity model file {fn!r} not found")
            layers, files = md.read_lhm_rmed(fn)
            d = ini.get("dir_rmed", "")
            rhomin = float(ini.get("rhomin", 1.0, "s"))
            cache = {}
            fields = []
            for f in files:
                p = d + "/" + f
                if p not in cache:
                    cache[p] = _rmed_field(g, dim, p)
                fields.append(cache[p])
            if all(f is None for f in fields):
                xi = np.zeros(shape, F32)
            else:
                xi = md.layer_xi(zc, layers, vcut, fields)
            out = md.vmodel_lhm(zc, layers, vcut, xi=xi, vmax_clip=vmax, rhomin=rhomin, munk=munk)
        return (*out, float(F32(layers[0].depth)))
    raise InputError(f"vmodel_type {vtype!r} is not supported by this host (uni, uni_rmed, lhm, lhm_rmed are)")


# ---------------------------------------------------------------------------
# sources
# ---------------------------------------------------------------------------
def _cell_mu(g, med, dim):
    def mu_at(x, y, z):
        i = fdtool.x2i(x, g.xbeg, g.dx)
        k = fdtool.x2i(z, g.zbeg, g.dz)
        if dim == "3d":
            j = fdtool.x2i(y, g.ybeg, g.dy)
            if -2 <= i <= g.nx + 3 and -2 <= j <= g.ny + 3 and -2 <= k <= g.nz + 3:
                return float(med.mu[j + 2, i + 2, k + 2])
        elif -2 <= i <= g.nx + 3 and -2 <= k <= g.nz + 3:
            return float(med.mu[i + 2, k + 2])
        return 0.0
    return mu_at


def sources(ini, g, dim, bench, med, bddep0, clon, clat, phi, nt):
    """(glut, bforce, hypocentre info) of source__setup."""
    stftype = ini.get("stftype", "kupper")
    if stftype == "scosine":
        stftype = "cosine"
    bf = ini.get("bf_mode", False) and not bench
    if bench:
        m = float(F32(1.0) / F32(math.sqrt(3.0)))
        src = [PointSource(0.0, 0.0, 5.0, 0.1, 2.0, 1e15, (m, m, m, 0.0, 0.0, 0.0))]
        info = {"evlo": clon, "evla": clat, "evdp": 5.0, "sx0": 0.0, "sy0": 0.0, "m6": src[0].m6, "otim": 0.0}
        stftype = "kupper"
    else:
        fn = ini.get("fn_stf", "")
        if not os.path.exists(fn):
            raise InputError(f"source file {fn!r} not found")
        fmt = ini.get("stf_format", "xym0ij")
        if bf:
            src, info = read_bodyforce_file(fn, fmt, phi, clon, clat, with_info=True)
        else:
            src, info = read_source_file(fn, fmt, phi, clon, clat, _cell_mu(g, med, dim), with_info=True)
    sdep = ini.get("sdep_fit", "asis")
    if sdep != "asis":
        if len(sdep) == 3 and sdep[:2] == "bd" and sdep[2].isdigit():
            depth = bddep0 if sdep == "bd0" else -9999.0
            src = [type(s)(**{**vars(s), "z": float(F32(depth))}) for s in src]
        else:
            raise InputError(f"unknown sdep_fit {sdep!r}")
    if dim == "3d":
        if bf:
            return None, body_force3(g, src, nt, stftype), info
        return stress_glut3(g, src, nt, stftype), None, info
    if dim == "sh":
        if bf:
            return None, body_force_sh(g, src, nt, stftype), info
        return stress_glut_sh(g, src, nt, stftype), None, info
    if bf:
        return None, body_force2(g, src, nt, stftype), info
    return stress_glut2(g, src, nt, stftype), None, info


# ---------------------------------------------------------------------------
# setup
# ---------------------------------------------------------------------------
def setup(inp, dim="3d"):
    ini = _ini(inp)
    _unsupported(ini, dim)
    title = ini.get("title", {"3d": "swpc3d", "psv": "swpc_psv", "sh": "swpc_sh"}[dim])
    nt = ini.get("nt", 1000)
    odir = ini.get("odir", "./out")
    g, bench, clon, clat, phi = _grid(ini, dim)
    rho, lam, mu, qp, qs, bddep0 = velocity_model(ini, g, dim, bench)
    fq_min = float(ini.get("fq_min", 0.05, "s"))
    fq_max = float(ini.get("fq_max", 5.0, "s"))
    fq_ref = float(ini.get("fq_ref", 1.0, "s"))
    if dim == "3d":
        med = md.medium_setup3(g, rho, lam, mu, qp, qs, fq_min, fq_max, fq_ref, benchmark=bench)
    else:
        med = md.medium_setup2(g, rho, lam, mu, qp, qs, fq_min, fq_max, fq_ref, benchmark=bench)
    glut, bforce, info = sources(ini, g, dim, bench, med, bddep0, clon, clat, phi, nt)
    tab = glut if glut is not None else bforce
    prof = pml.pml3(g, tab.fcut) if dim == "3d" else (pml.pml_sh(g, tab.fcut) if dim == "sh" else pml.pml2(g, tab.fcut))
    # report__setup: stability and wavelength conditions (m_report.f90:71-104)
    fmax = 2.0 * tab.fcut
    c = problem.stability3(g, med.vmax) if dim == "3d" else \
        med.vmax * g.dt / (1.0 / math.sqrt(1 / g.dx ** 2 + 1 / g.dz ** 2)) * 7.0 / 6.0
    dh = max(g.dx, getattr(g, "dy", g.dx), g.dz)
    r = med.vmin / fmax / dh if fmax > 0 else float("inf")
    if c > 1.0:
        raise InputError(f"stability condition is violated (c = {c:.3f} > 1): use a smaller dt")
    # stations
    sw_v = ini.get("sw_wav_v", False)
    sw_u = ini.get("sw_wav_u", False)
    ntdec_w = ini.get("ntdec_w", 10)
    stations, cells, lonlat = [], np.zeros((0, 3 if dim == "3d" else 2), np.int32), []
    if sw_v or sw_u:
        fn = ini.get("fn_stloc", "")
        if os.path.exists(fn):
            sts, ll = read_station_file(fn, ini.get("st_format", "xy"), clon, clat, phi, with_lonlat=True)
            if dim == "3d":
                cells, stations = station_cells3(g, med.kfs, med.kob, sts)
            else:
                cells, stations = station_cells2(g, med.kfs, med.kob, sts)
            keep = {id(s) for s in stations}
            lonlat = [l for s, l in zip(sts, ll) if id(s) in keep]
        else:
            print("[info] no station location file found", file=sys.stderr)
    # snapshots
    cfg = {"ntdec_s": ini.get("ntdec_s", 10), "fmt": ini.get("snp_format", "native"),
           "idec": ini.get("idec", 1), "jdec": ini.get("jdec", 1), "kdec": ini.get("kdec", 1)}
    if cfg["fmt"] not in ("native", "netcdf"):
        raise InputError(f"unknown snp_format {cfg['fmt']!r}")
    snaps = []
    if dim == "3d":
        cfg["z0_xy"] = float(ini.get("z0_xy", F32(max(min(10.0, g.zend), g.zbeg)), "s"))
        cfg["x0_yz"] = float(ini.get("x0_yz", F32(max(min(0.0, g.xend), g.xbeg)), "s"))
        cfg["y0_xz"] = float(ini.get("y0_xz", F32(max(min(0.0, g.yend), g.ybeg)), "s"))
        snaps = [(p, k) for (p, k) in SNAP3 if ini.get(f"{p}_{k}%sw", False)]
    else:
        # swpc_sh reads only xz_v%sw and xz_u%sw (src/swpc_sh/m_snap.f90:82-83)
        snaps = [("xz", k) for k in (("v", "u") if dim == "sh" else ("ps", "v", "u"))
                 if ini.get(f"xz_{k}%sw", False)]
    wav = {"v": sw_v, "u": sw_u, "format": ini.get("wav_format", "sac"), "prg": ini.get("ntdec_w_prg", 0)}
    if wav["format"] not in ("sac", "csf", "tar_st", "tar_node"):
        raise InputError(f"unknown wav_format {wav['format']!r}")
    return Setup(dim, ini, title, odir, nt, g, med, prof, glut, bforce, stations, cells, lonlat, ntdec_w,
                 ini.get("ntdec_r", 10), wav, snaps, cfg, info, clon, clat, phi, bddep0,
                 {"c": c, "r": r, "vmin": med.vmin, "vmax": med.vmax, "fmax": fmax})


# ---------------------------------------------------------------------------
# outputs
# ---------------------------------------------------------------------------
def _wav_meta(S, exedate, M0):
    info = S.meta
    mag = fdtool.moment_magnitude(M0) if S.glut is not None else sac.UNDEF_F
    return sac.WavMeta(title=S.title, dt=float(S.grid.dt), ntdec_w=S.ntdec_w,
                       ntw=0, tbeg=float(S.grid.tbeg), evlo=info["evlo"], evla=info["evla"], evdp=info["evdp"],
                       sx0=info["sx0"], sy0=info["sy0"], mag=float(mag), otim=info["otim"],
                       m6=info.get("m6") if S.glut is not None else None,
                       f3=info.get("f3") if S.bforce is not None else None,
                       clon=S.clon, clat=S.clat, phi=S.phi, exedate=exedate)


def write_waveforms(S, sim, meta):
    if not S.stations or not (S.wav["v"] or S.wav["u"]):
        return []
    vel, disp = sim.receivers(scaled=True)
    meta.ntw = sim.ntw
    if S.dim == "sh":   # one component (Vy, Uy)
        vel, disp = vel[:, :, None], disp[:, :, None]
    comp = {"3d": [0, 1, 2], "psv": [0, 1], "sh": [0]}[S.dim]
    v = vel[:, :, comp] if S.wav["v"] else None
    u = disp[:, :, comp].astype(F32) if S.wav["u"] else None
    fmt = S.wav["format"]
    if fmt == "sac":
        return sac.write_wav(S.odir, meta, S.stations, v, u, S.dim, S.stlonlat)
    if fmt == "csf":
        return [sac.write_csf(S.odir, meta, S.stations, v, u, S.dim, S.stlonlat)]
    return sac.write_tar(S.odir, meta, S.stations, v, u, S.dim, S.stlonlat,
                         per="station" if fmt == "tar_st" else "rank")


def run(inp, dim="3d", precision="fp32", log=None, exedate=None):
    """Set up and run the simulation of an input.inf; returns the Setup with
    report["wall"] and report["paths"] filled in."""
    from .solver import NumericalDivergence, Swpc3D, SwpcPSV, SwpcSH, DIVERGENCE_TOL
    log = log if log is not None else (lambda s: print(s, file=sys.stderr, flush=True))
    S = setup(inp, dim)
    g = S.grid
    os.makedirs(S.odir, exist_ok=True)
    exedate = int(time.time()) if exedate is None else int(exedate)
    rp = S.report
    log(f"  Grid Size               : {g.nx} x {getattr(g, 'ny', 1)} x {g.nz}")
    log(f"  Stability  Condition c  : {rp['c']:15.3f}  (c<1)")
    log(f"  Wavelength Condition r  : {rp['r']:15.3f}  (r>5-10)")
    log(f"  Minimum velocity        : {rp['vmin']:15.3f}  [km/s]")
    log(f"  Maximum velocity        : {rp['vmax']:15.3f}  [km/s]")
    log(f"  Maximum frequency       : {rp['fmax']:15.3f}  [Hz]")
    Sim = {"3d": Swpc3D, "psv": SwpcPSV, "sh": SwpcSH}[dim]
    cells = S.cells if len(S.cells) else None
    with Sim(g, S.medium, S.prof, S.glut, cells, S.ntdec_w, S.nt, precision=precision, bodyforce=S.bforce) as sim:
        meta = _wav_meta(S, exedate, sim.M0)
        ss = None
        if S.snaps:
            cfg = S.snap_cfg
            ss = snapshot.SnapshotSet(sim, S.odir, S.title, ntdec_s=cfg["ntdec_s"], nbuf=8, exedate=exedate,
                                      clon=S.clon, clat=S.clat, phi=S.phi, fmt=cfg["fmt"],
                                      ev={k: S.meta[k] for k in ("evlo", "evla", "evdp")} |
                                      {"evx": S.meta["sx0"], "evy": S.meta["sy0"]})
            if dim == "3d":
                sg = snapshot.snap_grid3(g, cfg["idec"], cfg["jdec"], cfg["kdec"], cfg["z0_xy"], cfg["x0_yz"],
                                         cfg["y0_xz"])
                for p, k in S.snaps:
                    ss.add3(sg, p, k, S.medium, topo=S.bddep0)
            else:
                sg = snapshot.snap_grid2(g, cfg["idec"], cfg["kdec"])
                for _, k in S.snaps:
                    (ss.add_sh if dim == "sh" else ss.add2)(sg, k, S.medium)
        ring = S.snap_cfg["ntdec_s"] * 8 if ss else S.nt
        prg = S.wav["prg"]
        t0 = time.perf_counter()
        it = 1
        while it <= S.nt:
            if it % S.ntdec_r == 0:
                vm = sim.kernel__vmax().astype(np.float64)
                amp = float(vm.max()) * sim.UC
                if not math.isfinite(amp) or amp > DIVERGENCE_TOL:
                    raise NumericalDivergence(f"numerical divergence detected at it={it} with max amp = {amp:.3e}")
                el = time.perf_counter() - t0
                per = el / (it - 1) if it > 1 else 0.0
                eta = int(per * (S.nt - it + 1))
                log(f"  it={it:07d},{per:6.3f} s/loop, eta {eta // 3600:03d}:{eta % 3600 // 60:02d}:{eta % 60:02d}, ("
                    + " ".join(f"{v:9.2E}" for v in vm * sim.UC * sim.M0) + " )")
            stop = min(S.nt + 1, (it // S.ntdec_r + 1) * S.ntdec_r, it + ring)
            if prg > 0:
                stop = min(stop, ((it - 1) // prg + 1) * prg + 1)
            sim.run(it, stop - it)
            if ss:
                ss.flush(stop - 1)
            # progressive waveform export at mod(it-1, ntdec_w_prg) = 0 (m_wav.f90:622-624)
            if prg > 0 and (stop - 1) % prg == 0 and stop - 1 < S.nt:
                write_waveforms(S, sim, meta)
            it = stop
        sim.sync()
        rp["wall"] = time.perf_counter() - t0
        rp["paths"] = write_waveforms(S, sim, meta)
        if ss:
            ss.close(S.nt)
            rp["paths"] += [s.file.path for s in ss.streams]
    log(f"  total time {rp['wall']:.3f} s for {S.nt} steps")
    return S


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("-i", default="./in/input.inf", help="parameter file (input.inf)")
    ap.add_argument("-d", "--dim", default="3d", choices=["3d", "psv", "sh"])
    ap.add_argument("-p", "--precision", default="fp32", choices=["fp32", "fp64", "fp32_s16", "fp32_r16"])
    a = ap.parse_args(argv)
    try:
        run(a.i, a.dim, a.precision)
    except InputError as e:
        print(f"[error] {e}", file=sys.stderr)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main())
