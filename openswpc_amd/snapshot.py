"""Snapshot output: decimated sections, their native file format, and the
streaming of device frames to files.

Restates the host side of ``snap__setup``/``snap__write``/
``snap__closefiles`` (src/swpc_3d/m_snap.f90:97-323, 838-908, 911-2185,
2202-2289; P-SV src/swpc_psv/m_snap.f90:60-160, 260-315, 419-650).  The
products themselves are computed on the GPU (``swpc3d_add_snapshot`` /
``swpc2d_add_snapshot``); this module decides which cells each section
samples, writes the file header and medium planes, and appends frames.

Native file (``snp_format = 'native'``), machine-endian stream:
  header (write_snp_header, m_snap.f90:838-884): 'STREAMIO', 'SWPC_3D '
  or 'SWPC_PSV', int32 6, title char(80), int32 exedate, coordinate
  char(2), snaptype char(2), int32 ns1, ns2, float32 beg1, beg2, ds1, ds2,
  dt*ntdec_s, int32 na1, na2, nmed, nsnp, float32 clon, clat, phi and
  three float32 words (1.0 in 3-D, 0.0 in P-SV);
  medium planes: rho, lam, mu (ns1 x ns2, first index fastest), plus the
  topography -bddep*1000 for xy/fs/ob sections (newfile_xy 417-470) -- the
  header's nmed counts 6 there (the netCDF layout), while the native file
  holds these 4 planes, which is what src/tools/read_snp.f90:434-453 reads;
  then nsnp planes per frame.

netCDF file (``snp_format = 'netcdf'``, classic format, written by
``openswpc_amd.cdf``), following write_nc_header / newfile_*_nc /
wbuf_nc / close_nc / output__put_maxval (m_snap.f90:475-741, 942-971,
2185-2344; P-SV src/swpc_psv/m_snap.f90:160-260, 317-390, 652-720):
dimensions (x1, x2, t unlimited); coordinate variables x1, x2, t; medium
rho, lambda, mu (plus topo, lon, lat on xy/fs/ob sections); one record
variable per product; max-V, max-H, max-A on fs/ob velocity and
displacement sections.  Record r holds the frame of it0 = r*ntdec_s + 1
at index it0/ntdec_s (so with ntdec_s = 1 the first record stays at the
fill value, as in the reference) and t = it0*dt.  Variables are declared
with the reference's dimension order (x1 fastest), attributes in its
order.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import cdf, fdtool, geomap

F32 = np.float32
KIND = {"ps": 0, "v": 1, "u": 2}
PLANES3 = ("yz", "xz", "xy", "fs", "ob")
UC3 = F32(1e-15)   # real(SP) UC of src/swpc_3d/m_global.f90:29
UC2 = F32(1e-12)   # src/swpc_psv/m_global.f90:29


def _ceil_div(a, b):
    return -((-a) // b)


@dataclass
class SnapGrid3:
    """Decimated snapshot grid (m_snap.f90:127-162)."""
    idec: int
    jdec: int
    kdec: int
    nxs: int
    nys: int
    nzs: int
    xsnp: np.ndarray
    ysnp: np.ndarray
    zsnp: np.ndarray
    i0_yz: int
    j0_xz: int
    k0_xy: int


def snap_grid3(g, idec=1, jdec=1, kdec=1, z0_xy=None, x0_yz=None, y0_xz=None):
    """Section sizes, coordinates and section positions; defaults as the
    reference's readini defaults (m_snap.f90:116-118)."""
    z0 = F32(max(min(10.0, g.zend), g.zbeg)) if z0_xy is None else F32(z0_xy)
    x0 = F32(max(min(0.0, g.xend), g.xbeg)) if x0_yz is None else F32(x0_yz)
    y0 = F32(max(min(0.0, g.yend), g.ybeg)) if y0_xz is None else F32(y0_xz)
    nxs = (g.nx + idec // 2) // idec
    nys = (g.ny + jdec // 2) // jdec
    nzs = (g.nz + kdec // 2) // kdec
    xs = np.array([fdtool.i2x(i * idec - idec // 2, g.xbeg, g.dx) for i in range(1, nxs + 1)], F32)
    ys = np.array([fdtool.i2x(j * jdec - jdec // 2, g.ybeg, g.dy) for j in range(1, nys + 1)], F32)
    zs = np.array([fdtool.i2x(k * kdec - kdec // 2, g.zbeg, g.dz) for k in range(1, nzs + 1)], F32)
    return SnapGrid3(idec, jdec, kdec, nxs, nys, nzs, xs, ys, zs,
                     int(fdtool.x2i(x0, g.xbeg, g.dx)), int(fdtool.x2i(y0, g.ybeg, g.dy)),
                     int(fdtool.x2i(z0, g.zbeg, g.dz)))


def cells3(sg: SnapGrid3, plane, kfs=None, kob=None):
    """(n1, n2, ijk) of a 3-D section: ijk[(a-1) + n1*(b-1)] = (i, j, k),
    1-based, with the first section index fastest as the reference's
    buffers.  kfs/kob are the (ny+6, nx+6) surface maps."""
    if plane == "yz":
        a, b = np.meshgrid(np.arange(1, sg.nys + 1), np.arange(1, sg.nzs + 1), indexing="xy")
        j = a * sg.jdec - sg.jdec // 2
        k = b * sg.kdec - sg.kdec // 2
        i = np.full_like(j, sg.i0_yz)
        n1, n2 = sg.nys, sg.nzs
    elif plane == "xz":
        a, b = np.meshgrid(np.arange(1, sg.nxs + 1), np.arange(1, sg.nzs + 1), indexing="xy")
        i = a * sg.idec - sg.idec // 2
        k = b * sg.kdec - sg.kdec // 2
        j = np.full_like(i, sg.j0_xz)
        n1, n2 = sg.nxs, sg.nzs
    elif plane in ("xy", "fs", "ob"):
        a, b = np.meshgrid(np.arange(1, sg.nxs + 1), np.arange(1, sg.nys + 1), indexing="xy")
        i = a * sg.idec - sg.idec // 2
        j = b * sg.jdec - sg.jdec // 2
        if plane == "xy":
            k = np.full_like(i, sg.k0_xy)
        else:
            m = kfs if plane == "fs" else kob
            if m is None:
                raise ValueError(f"{plane} sections need the {'kfs' if plane == 'fs' else 'kob'} map")
            k = np.asarray(m)[j + 2, i + 2] + 1
        n1, n2 = sg.nxs, sg.nys
    else:
        raise ValueError(f"unknown section {plane!r}")
    ijk = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.int32)
    return n1, n2, ijk


@dataclass
class SnapGrid2:
    idec: int
    kdec: int
    nxs: int
    nzs: int
    xsnp: np.ndarray
    zsnp: np.ndarray


def snap_grid2(g, idec=1, kdec=1):
    """src/swpc_psv/m_snap.f90:90-106."""
    nxs = (g.nx + idec // 2) // idec
    nzs = (g.nz + kdec // 2) // kdec
    xs = np.array([fdtool.i2x(i * idec - idec // 2, g.xbeg, g.dx) for i in range(1, nxs + 1)], F32)
    zs = np.array([fdtool.i2x(k * kdec - kdec // 2, g.zbeg, g.dz) for k in range(1, nzs + 1)], F32)
    return SnapGrid2(idec, kdec, nxs, nzs, xs, zs)


def cells2(sg: SnapGrid2):
    a, b = np.meshgrid(np.arange(1, sg.nxs + 1), np.arange(1, sg.nzs + 1), indexing="xy")
    i = a * sg.idec - sg.idec // 2
    k = b * sg.kdec - sg.kdec // 2
    return sg.nxs, sg.nzs, np.stack([i.ravel(), k.ravel()], axis=1).astype(np.int32)


def _c(s, n):
    return str(s).encode("ascii", "replace")[:n].ljust(n, b" ")


def header_bytes(code, title, exedate, coordinate, snaptype, xs1, xs2, dt, ntdec_s, na1, na2, nmed, nsnp,
                 clon=0.0, clat=0.0, phi=0.0, tail=1.0):
    """write_snp_header (m_snap.f90:838-884; P-SV 272-315)."""
    xs1 = np.asarray(xs1, F32)
    xs2 = np.asarray(xs2, F32)
    d1 = F32(xs1[1] - xs1[0]) if xs1.size > 1 else F32(0)
    d2 = F32(xs2[1] - xs2[0]) if xs2.size > 1 else F32(0)
    parts = [b"STREAMIO", _c(code, 8), np.int32(6).tobytes(), _c(title, 80), np.int32(exedate).tobytes(),
             _c(coordinate, 2), _c(snaptype, 2), np.array([xs1.size, xs2.size], np.int32).tobytes(),
             np.array([xs1[0], xs2[0], d1, d2, F32(F32(dt) * F32(ntdec_s))], F32).tobytes(),
             np.array([na1, na2, nmed, nsnp], np.int32).tobytes(),
             np.array([clon, clat, phi, tail, tail, tail], F32).tobytes()]
    return b"".join(parts)


def _plane(a):
    """(n1, n2) array -> bytes with the first index fastest."""
    return np.asarray(a, F32).T.tobytes()


class SnapFile:
    """One native snapshot file: header and medium at creation, then one
    frame of nsnp planes per append."""

    def __init__(self, path, header, medium_planes):
        self.path = path
        self.f = open(path, "wb")
        self.f.write(header)
        for p in medium_planes:
            self.f.write(_plane(p))
        self.frames = 0

    def append(self, planes):
        for p in planes:
            self.f.write(_plane(p))
        self.frames += 1

    def close(self):
        if self.f:
            self.f.close()
            self.f = None


def read_snp(path):
    """Read a native snapshot file: (header dict, medium planes, frames
    [nframe, nsnp, n1, n2])."""
    raw = open(path, "rb").read()
    o = 0

    def take(n):
        nonlocal o
        b = raw[o:o + n]
        o += n
        return b

    h = {"bintype": take(8).decode(), "codetype": take(8).decode(),
         "hdrver": int(np.frombuffer(take(4), np.int32)[0]), "title": take(80).decode().rstrip(),
         "exedate": int(np.frombuffer(take(4), np.int32)[0]), "coordinate": take(2).decode(),
         "datatype": take(2).decode()}
    h["ns1"], h["ns2"] = (int(v) for v in np.frombuffer(take(8), np.int32))
    h["beg1"], h["beg2"], h["ds1"], h["ds2"], h["dt"] = (float(v) for v in np.frombuffer(take(20), F32))
    h["na1"], h["na2"], h["nmed"], h["nsnp"] = (int(v) for v in np.frombuffer(take(16), np.int32))
    h["clon"], h["clat"], h["phi"], *h["tail"] = (float(v) for v in np.frombuffer(take(24), F32))
    n1, n2 = h["ns1"], h["ns2"]
    nmed_file = 4 if h["coordinate"] in ("xy", "fs", "ob") else 3
    psz = 4 * n1 * n2

    def plane():
        return np.frombuffer(take(psz), F32).reshape(n2, n1).T.copy()

    med = [plane() for _ in range(nmed_file)]
    frames = []
    while len(raw) - o >= psz * h["nsnp"]:
        frames.append(np.stack([plane() for _ in range(h["nsnp"])]))
    return h, med, np.array(frames, F32).reshape(-1, h["nsnp"], n1, n2)


VNAMES3 = {"ps": (("div", "rot_x", "rot_y", "rot_z"), ("1/s",) * 4),
           "v": (("Vx", "Vy", "Vz"), ("m/s",) * 3), "u": (("Ux", "Uy", "Uz"), ("m",) * 3)}
VNAMES2 = {"ps": (("divergence", "rotation"), ("1/s",) * 2),
           "v": (("Vx", "Vz"), ("m/s",) * 2), "u": (("Ux", "Uz"), ("m",) * 2)}
VNAMES_SH = {"v": (("Vy",), ("m/s",)), "u": (("Uy",), ("m",))}
PEAK_NAMES = (("max-V", "Maximum amplitude of the vertical component"),
              ("max-H", "Maximum amplitude of the horizontal components"),
              ("max-A", "Maximum amplitude of the vector motion"))


class NcSnapFile:
    """One netCDF snapshot file.  ``head`` carries the global attribute
    values: title, exedate, na1, na2, ds1, ds2, dt_s (= dt*ntdec_s), evlo,
    evla, evdp, evx, evy, clon, clat, phi."""

    def __init__(self, path, code, plane, snaptype, kind, xs1, xs2, medium, head, dt, ntdec_s,
                 lonlat=None, peaks=False):
        # the P-SV and SH headers share one attribute order
        psv = code in ("SWPC_PSV", "SWPC_SH ")
        horiz = plane in ("xy", "fs", "ob")
        n1, n2 = len(xs1), len(xs2)
        names = {"yz": ("y", "z"), "xz": ("x", "z")}.get(plane, ("x", "y"))
        vn, vu = {"SWPC_PSV": VNAMES2, "SWPC_SH ": VNAMES_SH}.get(code, VNAMES3)[kind]
        self.path, self.nvar, self.n1, self.n2 = path, len(vn), n1, n2
        self.dt, self.ntdec_s, self.peaks = F32(dt), int(ntdec_s), peaks
        self.vnames = vn
        self.vmin = [F32(0)] * len(vn)
        self.vmax = [F32(0)] * len(vn)
        w = cdf.Writer(path, 1)
        w.dim(names[0], n1)
        w.dim(names[1], n2)
        w.dim("t", None)
        rng = lambda a: np.array([np.min(a), np.max(a)], F32)   # noqa: E731
        x1 = {"long_name": names[0]}
        x2 = {"long_name": names[1]}
        if horiz:
            x1["standard_name"] = "projection_x_coordinate"
            x2["standard_name"] = "projection_y_coordinate"
        for a in (x1, x2):
            a["units"] = "km"
        x1["actual_range"] = np.array([xs1[0], xs1[-1]], F32)
        x2["actual_range"] = np.array([xs2[0], xs2[-1]], F32)
        w.var(names[0], "f4", [names[0]], x1)
        w.var(names[1], "f4", [names[1]], x2)
        w.var("t", "f4", ["t"], {"long_name": "t", "units": "s"})
        dims2 = [names[1], names[0]]   # Fortran (x1, x2): x1 fastest
        munits = ("10^3 kg/cm^3", "10^9 Pa", "10^9 Pa")
        for (nm, arr), un in zip(zip(("rho", "lambda", "mu"), medium[:3]), munits):
            a = {"coordinates": "lat lon"} if horiz else {}
            a.update({"long_name": nm, "units": un, "actual_range": rng(arr)})
            w.var(nm, "f4", dims2, a)
        if horiz:
            w.var("topo", "f4", dims2, {"long_name": "Topography", "units": "m", "coordinates": "lat lon",
                                        "actual_range": rng(medium[3])})
            w.var("lon", "f4", dims2, {"long_name": "Longitude", "units": "degrees_east",
                                       "actual_range": rng(lonlat[0])})
            w.var("lat", "f4", dims2, {"long_name": "Latitude", "units": "degrees_north",
                                       "actual_range": rng(lonlat[1])})
        for nm, un in zip(vn, vu):
            a = {"long_name": nm, "units": un}
            if horiz:
                a["coordinates"] = "lat lon"
            a["actual_range"] = np.zeros(2, F32)
            w.var(nm, "f4", ["t"] + dims2, a)
        if peaks:
            for nm, ln in PEAK_NAMES:
                w.var(nm, "f4", dims2, {"long_name": ln, "coordinates": "lat lon",
                                        "units": "m/s" if kind == "v" else "m", "actual_range": np.zeros(2, F32)})
        h = head
        g = [("generated_by", "SWPC")]
        common = [("ns1", np.int32(n1)), ("ns2", np.int32(n2)), ("beg1", F32(xs1[0])), ("beg2", F32(xs2[0])),
                  ("na1", np.int32(h["na1"])), ("na2", np.int32(h["na2"])), ("ds1", F32(h["ds1"])),
                  ("ds2", F32(h["ds2"])), ("nmed", np.int32(6 if horiz else 3)), ("nsnp", np.int32(len(vn))),
                  ("coordinate", plane), ("datatype", snaptype), ("dt", F32(h["dt_s"])),
                  ("evlo", F32(h["evlo"])), ("evla", F32(h["evla"])), ("evdp", F32(h["evdp"])),
                  ("evx", F32(h["evx"])), ("evy", F32(h["evy"]))]
        ident = [("title", h["title"].strip()), ("exedate", np.int32(h["exedate"])), ("hdrver", np.int32(6)),
                 ("codetype", code)]
        if psv:
            g += ident + common
        else:
            g += [ident[3], ident[2], ident[0], ident[1]] + common
        g += [("clon", F32(h["clon"])), ("clat", F32(h["clat"])), ("phi", F32(h["phi"]))]
        for k, v in g:
            w.gatt(k, v)
        w.enddef()
        w.put(names[0], np.asarray(xs1, F32))
        w.put(names[1], np.asarray(xs2, F32))
        for nm, arr in zip(("rho", "lambda", "mu"), medium[:3]):
            w.put(nm, np.asarray(arr, F32).T)
        if horiz:
            w.put("topo", np.asarray(medium[3], F32).T)
            w.put("lon", np.asarray(lonlat[0], F32).T)
            w.put("lat", np.asarray(lonlat[1], F32).T)
        self.w = w
        self.frames = 0

    def append(self, planes, frame):
        """Frame number frame (it0 = frame*ntdec_s + 1); planes (n1, n2)."""
        it0 = frame * self.ntdec_s + 1
        rec = it0 // self.ntdec_s
        for v, p in enumerate(planes):
            p = np.asarray(p, F32)
            self.w.put_record(self.vnames[v], rec, p.T)
            self.w.put_record("t", rec, np.array([F32(it0) * self.dt], F32))
            self.vmax[v] = max(self.vmax[v], F32(p.max()))
            self.vmin[v] = min(self.vmin[v], F32(p.min()))
        self.frames += 1

    def close(self, peaks=None):
        if self.w is None:
            return
        for v, nm in enumerate(self.vnames):
            self.w.att(nm, "actual_range", np.array([self.vmin[v], self.vmax[v]], F32))
        if self.peaks and peaks is not None:
            for (nm, _), pk in zip(PEAK_NAMES, peaks):
                pk = np.asarray(pk, F32)
                self.w.att(nm, "actual_range", np.array([pk.min(), pk.max()], F32))
                self.w.put(nm, pk.T)
        self.w.close()
        self.w = None


def read_nc_snp(path):
    """(global attributes, {variable: data}) of a netCDF snapshot file; 2-D
    variables come back as (n1, n2) planes and record variables as
    (nrec, n1, n2)."""
    ds = cdf.read(path)
    out = {}
    for nm, v in ds.variables.items():
        d = v.data()
        if d.ndim == 2:
            d = d.T
        elif d.ndim == 3:
            d = d.transpose(0, 2, 1)
        out[nm] = d
    return ds.gatts, out


@dataclass
class _Stream:
    sid: int
    plane: str
    kind: str
    n1: int
    n2: int
    nvar: int
    file: object
    written: int = 0
    peaks: bool = False


@dataclass
class SnapshotSet:
    """Snapshot streams of one solver, written under odir as native files
    (``<odir>/<title>.3d.<sec>.<kind>.snp``, ``.psv.xz.<kind>.snp``,
    ``.sh.xz.<kind>.snp``) or,
    with fmt = 'netcdf', as classic netCDF files (``.nc``).

    ``run(it0, n)`` advances the solver in chunks that fit the device ring
    and appends every finished frame to its file; ``close(nt)`` ends the
    files (in netCDF mode after the closing update of step nt+1 and with
    the max-V/H/A fields of fs/ob velocity and displacement sections).
    ``ev`` holds the hypocentre attributes of the netCDF header (evlo,
    evla, evdp, evx, evy)."""
    sim: object
    odir: str
    title: str
    ntdec_s: int = 10
    nbuf: int = 8
    exedate: int = 0
    clon: float = 0.0
    clat: float = 0.0
    phi: float = 0.0
    fmt: str = "native"
    ev: dict = field(default_factory=dict)
    streams: list = field(default_factory=list)

    def _head(self, na1, na2, ds1, ds2):
        h = {"title": self.title, "exedate": self.exedate, "na1": na1, "na2": na2, "ds1": ds1, "ds2": ds2,
             "dt_s": F32(F32(self.sim.g.dt) * F32(self.ntdec_s)), "clon": self.clon, "clat": self.clat,
             "phi": self.phi}
        for k in ("evlo", "evla", "evdp", "evx", "evy"):
            h[k] = self.ev.get(k, 0.0)
        return h

    def add3(self, sg: SnapGrid3, plane, kind, medium, topo=0.0):
        g = self.sim.g
        n1, n2, ijk = cells3(sg, plane, medium.kfs, medium.kob)
        sid = self.sim.add_snapshot(KIND[kind], ijk, self.ntdec_s, self.nbuf, F32(getattr(self.sim, "UC", UC3)),
                                    F32(self.sim.M0), F32(g.dt))
        ii, jj, kk = ijk[:, 0] + 2, ijk[:, 1] + 2, ijk[:, 2] + 2
        med = [np.asarray(a)[jj, ii, kk].reshape(n2, n1).T for a in (medium.rho, medium.lam, medium.mu)]
        if plane in ("xy", "fs", "ob"):
            t = np.broadcast_to(np.asarray(topo, F32), (g.ny + 6, g.nx + 6))
            med.append((-t[ijk[:, 1] + 2, ijk[:, 0] + 2] * F32(1000)).astype(F32).reshape(n2, n1).T)
        if plane == "yz":
            xs1, xs2, na1, na2 = sg.ysnp, sg.zsnp, g.na // sg.jdec, g.na // sg.kdec
        elif plane == "xz":
            xs1, xs2, na1, na2 = sg.xsnp, sg.zsnp, g.na // sg.idec, g.na // sg.kdec
        else:
            xs1, xs2, na1, na2 = sg.xsnp, sg.ysnp, g.na // sg.idec, g.na // sg.jdec
        nvar = 4 if kind == "ps" else 3
        nmed = 6 if plane in ("xy", "fs", "ob") else 3
        snaptype = "ps" if kind == "ps" else kind + "3"
        if self.fmt == "netcdf":
            # ds1/ds2 as newfile_*_nc sets them (m_snap.f90:493-494, 565-566, 761-762)
            if plane == "yz":
                ds1, ds2 = F32(sg.jdec * F32(g.dy)), F32(sg.kdec * F32(g.dz))
            elif plane == "xz":
                ds1, ds2 = F32(sg.idec * float(g.dx)), F32(sg.kdec * float(g.dz))
            else:
                ds1, ds2 = F32(sg.idec * float(g.dx)), F32(sg.jdec * float(g.dy))
            lonlat = None
            if plane in ("xy", "fs", "ob"):
                xa, yb = np.meshgrid(np.asarray(xs1, F32), np.asarray(xs2, F32), indexing="ij")
                lonlat = geomap.c2g(xa, yb, self.clon, self.clat, self.phi)
            peaks = plane in ("fs", "ob") and kind in ("v", "u")
            if peaks:
                self.sim.snapshot_peaks(sid)
            path = os.path.join(self.odir, f"{self.title.strip()}.3d.{plane}.{kind}.nc")
            f = NcSnapFile(path, "SWPC_3D ", plane, snaptype, kind, xs1, xs2, med, self._head(na1, na2, ds1, ds2),
                           g.dt, self.ntdec_s, lonlat, peaks)
            self.streams.append(_Stream(sid, plane, kind, n1, n2, nvar, f, peaks=peaks))
            return path
        hdr = header_bytes("SWPC_3D ", self.title, self.exedate, plane, snaptype,
                           xs1, xs2, g.dt, self.ntdec_s, na1, na2, nmed, nvar, self.clon, self.clat, self.phi, 1.0)
        path = os.path.join(self.odir, f"{self.title.strip()}.3d.{plane}.{kind}.snp")
        self.streams.append(_Stream(sid, plane, kind, n1, n2, nvar, SnapFile(path, hdr, med)))
        return path

    def add2(self, sg: SnapGrid2, kind, medium):
        g = self.sim.g
        n1, n2, ik = cells2(sg)
        sid = self.sim.add_snapshot(KIND[kind], ik, self.ntdec_s, self.nbuf, F32(getattr(self.sim, "UC", UC2)),
                                    F32(self.sim.M0), F32(g.dt))
        ii, kk = ik[:, 0] + 2, ik[:, 1] + 2
        med = [np.asarray(a)[ii, kk].reshape(n2, n1).T for a in (medium.rho, medium.lam, medium.mu)]
        if self.fmt == "netcdf":
            ds1, ds2 = F32(sg.idec * float(g.dx)), F32(sg.kdec * float(g.dz))
            path = os.path.join(self.odir, f"{self.title.strip()}.psv.xz.{kind}.nc")
            f = NcSnapFile(path, "SWPC_PSV", "xz", "ps" if kind == "ps" else kind + "2", kind, sg.xsnp, sg.zsnp,
                           med, self._head(g.na // sg.idec, g.na // sg.kdec, ds1, ds2), g.dt, self.ntdec_s)
            self.streams.append(_Stream(sid, "xz", kind, n1, n2, 2, f))
            return path
        hdr = header_bytes("SWPC_PSV", self.title, self.exedate, "xz", "ps" if kind == "ps" else kind + "2",
                           sg.xsnp, sg.zsnp, g.dt, self.ntdec_s, g.na // sg.idec, g.na // sg.kdec, 3, 2,
                           self.clon, self.clat, self.phi, 0.0)
        path = os.path.join(self.odir, f"{self.title.strip()}.psv.xz.{kind}.snp")
        self.streams.append(_Stream(sid, "xz", kind, n1, n2, 2, SnapFile(path, hdr, med)))
        return path

    def add_sh(self, sg: SnapGrid2, kind, medium):
        """An SH section (src/swpc_sh/m_snap.f90:73-143): 'v' (Vy) or 'u'
        (Uy), one plane per frame; snaptype 'vy'/'uy', the header's three
        trailing words -12345.0."""
        if kind not in ("v", "u"):
            raise ValueError(f"SH snapshots are 'v' or 'u', not {kind!r}")
        g = self.sim.g
        n1, n2, ik = cells2(sg)
        sid = self.sim.add_snapshot(KIND[kind], ik, self.ntdec_s, self.nbuf, F32(getattr(self.sim, "UC", UC2)),
                                    F32(self.sim.M0), F32(g.dt))
        ii, kk = ik[:, 0] + 2, ik[:, 1] + 2
        med = [np.asarray(a)[ii, kk].reshape(n2, n1).T for a in (medium.rho, medium.lam, medium.mu)]
        snaptype = kind + "y"
        if self.fmt == "netcdf":
            ds1, ds2 = F32(sg.idec * float(g.dx)), F32(sg.kdec * float(g.dz))
            path = os.path.join(self.odir, f"{self.title.strip()}.sh.xz.{kind}.nc")
            f = NcSnapFile(path, "SWPC_SH ", "xz", snaptype, kind, sg.xsnp, sg.zsnp, med,
                           self._head(g.na // sg.idec, g.na // sg.kdec, ds1, ds2), g.dt, self.ntdec_s)
            self.streams.append(_Stream(sid, "xz", kind, n1, n2, 1, f))
            return path
        hdr = header_bytes("SWPC_SH ", self.title, self.exedate, "xz", snaptype, sg.xsnp, sg.zsnp, g.dt,
                           self.ntdec_s, g.na // sg.idec, g.na // sg.kdec, 3, 1, self.clon, self.clat, self.phi,
                           -12345.0)
        path = os.path.join(self.odir, f"{self.title.strip()}.sh.xz.{kind}.snp")
        self.streams.append(_Stream(sid, "xz", kind, n1, n2, 1, SnapFile(path, hdr, med)))
        return path

    def flush(self, it_done):
        """Append frames produced by steps 1..it_done."""
        for s in self.streams:
            nf = (it_done - 1) // self.ntdec_s + 1 if it_done >= 1 else 0
            while s.written < nf:
                fr = self.sim.get_snapshot(s.sid, s.written, s.nvar, s.n1 * s.n2)
                planes = [fr[v].reshape(s.n2, s.n1).T for v in range(s.nvar)]
                if isinstance(s.file, NcSnapFile):
                    s.file.append(planes, s.written)
                else:
                    s.file.append(planes)
                s.written += 1

    def run(self, it0, n):
        chunk = self.ntdec_s * self.nbuf
        it = it0
        while it < it0 + n:
            m = min(chunk, it0 + n - it)
            self.sim.run(it, m)
            it += m
            self.flush(it - 1)

    def close(self, nt=None):
        """Close every file; in netCDF mode nt (the last step) is needed for
        the closing update at nt+1 that completes the maxima."""
        if self.fmt == "netcdf" and self.streams:
            if nt is None:
                raise ValueError("netCDF snapshots need nt at close")
            self.flush(nt)
            if any(s.peaks for s in self.streams):   # only the maxima observe the closing update
                self.sim.finish_snapshots(nt + 1)
        for s in self.streams:
            if isinstance(s.file, NcSnapFile):
                s.file.close(self.sim.get_peaks(s.sid, s.n1 * s.n2).reshape(3, s.n2, s.n1).transpose(0, 2, 1)
                             if s.peaks else None)
            else:
                s.file.close()
