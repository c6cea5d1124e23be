"""SAC waveform files and the receiver-trace export.

``SacHeader``/``write_sac``/``read_sac`` restate the binary SAC layout that
the reference writes (src/shared/m_sac.f90: ``sac__init`` 452-560,
``wsac_s`` 267-311, ``sac__whdr`` 314-449): 70 float32 words, 35 int32
words (header words 71-105), 5 four-byte logicals (106-110) and 48
four-character words (111-158), in machine byte order, followed by
``npts`` float32 samples.  Words the reference never writes keep the SAC
"undefined" values (-12345.0, -12345, false, '-12345').

``write_wav`` follows ``wav__write``/``export_wav__sac``
(src/swpc_3d/m_wav.f90:661-811) and the header set-up of
``set_sac_header``/``initialize_sac_header`` (m_wav.f90:276-398): one file
per station and component, ``<odir>/wav/<title>.3d.<stnm>.<cmp>.sac``
(``.psv.`` for the P-SV solver).
"""
from __future__ import annotations

import math
import os
import sys
import time
from dataclasses import dataclass, field

import numpy as np

F32 = np.float32
UNDEF_F = -12345.0
UNDEF_I = -12345
UNDEF_K = "-12345"

# float words 1..70 (1-based SAC word number) that sac__whdr fills
# (m_sac.f90:342-389)
REAL_WORDS = {
    "delta": 1, "depmin": 2, "depmax": 3, "b": 6, "e": 7, "o": 8, "a": 9,
    **{f"t{i}": 11 + i for i in range(10)},
    "stla": 32, "stlo": 33, "stel": 34, "stdp": 35, "evla": 36, "evlo": 37, "evel": 38, "evdp": 39,
    "mag": 40, **{f"user{i}": 41 + i for i in range(10)},
    "dist": 51, "az": 52, "baz": 53, "gcarc": 54, "depmen": 57, "cmpaz": 58, "cmpinc": 59,
    **{f"user{i}": 54 + i for i in range(10, 17)},
}
# integer words 71..105 (m_sac.f90:391-409)
INT_WORDS = {
    "nzyear": 71, "nzjday": 72, "nzhour": 73, "nzmin": 74, "nzsec": 75, "nzmsec": 76, "nvhdr": 77,
    "npts": 80, "iftype": 86, "idep": 87, "ievtyp": 93, **{f"iuser{i}": 98 + i for i in range(8)},
}
# logical words 106..110 (m_sac.f90:411-416)
LOGICAL_WORDS = {"leven": 106, "lpspol": 107, "lovrok": 108, "lcalda": 109, "luser0": 110}
# character fields: (first word, length in characters) (m_sac.f90:418-441)
CHAR_WORDS = {
    "kstnm": (111, 8), "kevnm": (113, 16), "khole": (117, 8), "ko": (119, 8), "ka": (121, 8),
    **{f"kt{i}": (123 + 2 * i, 8) for i in range(10)},
    "kf": (143, 8), "kuser0": (145, 8), "kuser1": (147, 8), "kuser2": (149, 8), "kcmpnm": (151, 8),
    "knetwk": (153, 8), "kdatrd": (155, 8), "kinst": (157, 8),
}
# SAC idep codes used by the reference (m_wav.f90:293, 304, 319)
IDEP_VEL, IDEP_DISP, IDEP_OTHER = 7, 6, 5


def _init_values():
    """The state ``sac__init`` leaves (m_sac.f90:452-560)."""
    v = {k: UNDEF_F for k in REAL_WORDS}
    v.update({k: UNDEF_I for k in INT_WORDS})
    v.update(nvhdr=6, iftype=1)
    v.update(leven=True, lpspol=False, lovrok=True, lcalda=False, luser0=False)
    v.update({k: UNDEF_K for k in CHAR_WORDS})
    return v


@dataclass
class SacHeader:
    """SAC header as a name -> value map holding the fields ``sac__whdr``
    writes; every field starts from the ``sac__init`` state."""
    values: dict = field(default_factory=_init_values)

    def __getattr__(self, name):
        vals = self.__dict__.get("values")
        if vals is not None and name in vals:
            return vals[name]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name == "values":
            object.__setattr__(self, name, value)
        elif name in self.values:
            self.values[name] = value
        else:
            raise AttributeError(f"unknown SAC header field {name!r}")

    def copy(self):
        return SacHeader(dict(self.values))

    def pack(self) -> bytes:
        """The 632-byte header block exactly as ``sac__whdr`` writes it."""
        f = np.full(70, UNDEF_F, F32)
        i = np.full(35, UNDEF_I, np.int32)
        lg = np.zeros(5, np.int32)
        a = bytearray(b"".join([b"-123", b"45  "] * 24))
        v = self.values
        # delta is truncated to 1e-7: real(int(delta*1d7))/1e7 in single precision
        d = float(v["delta"])
        f[0] = F32(F32(math.trunc(d * 1e7)) / F32(1e7))
        for k, w in REAL_WORDS.items():
            if w != 1:
                f[w - 1] = F32(v[k])
        for k, w in INT_WORDS.items():
            i[w - 71] = int(v[k])
        for k, w in LOGICAL_WORDS.items():
            lg[w - 106] = 1 if v[k] else 0
        for k, (w, n) in CHAR_WORDS.items():
            s = str(v[k]).encode("ascii", "replace")[:n].ljust(n, b" ")
            o = (w - 111) * 4
            a[o:o + n] = s
        return f.tobytes() + i.tobytes() + lg.tobytes() + bytes(a)


def write_sac(path, hdr: SacHeader, data):
    """Write one SAC file (``wsac_s``, m_sac.f90:267-311): the header block
    then the first ``npts`` samples as float32.  Existing files are
    overwritten (the reference's wav export passes overwrite = .true.)."""
    d = np.asarray(data, F32).ravel()
    n = int(hdr.npts)
    if n < 0 or n > d.size:
        raise ValueError(f"npts {n} outside the {d.size} samples given")
    with open(path, "wb") as fo:
        fo.write(hdr.pack())
        fo.write(d[:n].tobytes())


def read_sac(path):
    """Read a machine-endian SAC file: (SacHeader, float32 samples).  Byte
    order is detected from the header version word (77 = 6)."""
    raw = open(path, "rb").read()
    if len(raw) < 632:
        raise ValueError(f"{path}: shorter than a SAC header")
    bo = "<" if np.frombuffer(raw[304:308], "<i4")[0] == 6 else ">"
    f = np.frombuffer(raw[:280], bo + "f4")
    i = np.frombuffer(raw[280:420], bo + "i4")
    lg = np.frombuffer(raw[420:440], bo + "i4")
    a = raw[440:632]
    h = SacHeader()
    for k, w in REAL_WORDS.items():
        h.values[k] = float(f[w - 1])
    for k, w in INT_WORDS.items():
        h.values[k] = int(i[w - 71])
    for k, w in LOGICAL_WORDS.items():
        h.values[k] = bool(lg[w - 106])
    for k, (w, n) in CHAR_WORDS.items():
        o = (w - 111) * 4
        h.values[k] = a[o:o + n].decode("ascii", "replace").rstrip()
    npts = int(h.npts)
    data = np.frombuffer(raw[632:632 + 4 * npts], bo + "f4").astype(F32)
    return h, data


def _ymd2jul(y, m, d):
    """Day of year (m_daytim.f90:57-101)."""
    leap = (y % 4 == 0 and y % 100 != 0) or y % 400 == 0
    dom = [31, 29 if leap else 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31]
    return sum(dom[:m - 1]) + d


@dataclass
class WavMeta:
    """Run-level header values (``initialize_sac_header``,
    m_wav.f90:340-398).  Coordinates in km, angles in degrees."""
    title: str
    dt: float
    ntdec_w: int
    ntw: int
    tbeg: float = 0.0
    evlo: float = UNDEF_F
    evla: float = UNDEF_F
    evdp: float = UNDEF_F          # km
    sx0: float = 0.0               # x, y of the first source (km)
    sy0: float = 0.0
    mag: float = UNDEF_F
    otim: float = UNDEF_F          # origin time = first source's rupture start
    m6: tuple = None               # (mxx0, myy0, mzz0, myz0, mxz0, mxy0) of the first source
    f3: tuple = None               # (fx0, fy0, fz0) in body-force mode
    clon: float = UNDEF_F
    clat: float = UNDEF_F
    phi: float = 0.0
    exedate: int = None            # seconds since the epoch; None = now


def station_header(meta: WavMeta, stnm, xst, yst, zst, stlo=UNDEF_F, stla=UNDEF_F, dim="3d"):
    """Per-station header (``initialize_sac_header``, m_wav.f90:340-398;
    P-SV src/swpc_psv/m_wav.f90:640-680, where yst = sy0 makes the
    distance |x - sx0| and the azimuths 0 or 180 degrees)."""
    h = SacHeader()
    h.evlo, h.evla, h.evdp = meta.evlo, meta.evla, meta.evdp
    h.b = meta.tbeg
    h.delta = meta.ntdec_w * meta.dt
    h.npts = meta.ntw
    h.mag = meta.mag
    if dim == "3d":
        if meta.f3 is not None:
            h.user0, h.user1, h.user2 = meta.f3
        elif meta.m6 is not None:
            h.user0, h.user1, h.user2, h.user3, h.user4, h.user5 = meta.m6
    elif dim == "sh":
        pass   # the SH header leaves user0..5 undefined (src/swpc_sh/m_source.f90:458-463, 542-544)
    else:  # P-SV keeps the in-plane components only
        if meta.f3 is not None:
            h.user0, h.user2 = meta.f3[0], meta.f3[2]
        elif meta.m6 is not None:
            h.user0, h.user2, h.user4 = meta.m6[0], meta.m6[2], meta.m6[4]
    h.user6, h.user7, h.user8 = meta.clon, meta.clat, meta.phi
    h.o = meta.otim
    ex = int(time.time()) if meta.exedate is None else int(meta.exedate)
    lt = time.localtime(ex)
    h.nzyear, h.nzhour, h.nzmin, h.nzsec = lt.tm_year, lt.tm_hour, lt.tm_min, lt.tm_sec
    h.nzjday = _ymd2jul(lt.tm_year, lt.tm_mon, lt.tm_mday)
    h.nzmsec = 0
    h.kevnm = meta.title[:16].strip()
    h.kstnm = stnm.strip()
    h.stlo, h.stla = stlo, stla
    h.stdp = zst * 1000.0
    h.lcalda = False
    # distance and azimuths in single precision (x1, y1 are real(SP))
    dx, dy = F32(F32(xst) - F32(meta.sx0)), F32(F32(yst) - F32(meta.sy0))
    h.dist = float(np.sqrt(F32(dx * dx + dy * dy)))
    h.az = math.degrees(float(np.arctan2(dy, dx)))
    h.baz = math.degrees(float(np.arctan2(-dy, -dx)))
    return h


# component names, idep, cmpinc, cmpaz offset (m_wav.f90:285-334)
COMPONENTS_3D = {"v": [("Vx", 90.0, 0.0), ("Vy", 90.0, 90.0), ("Vz", 90.0, None)],
                 "u": [("Ux", 90.0, 0.0), ("Uy", 90.0, 90.0), ("Uz", 90.0, None)]}
COMPONENTS_PSV = {"v": [("Vx", 90.0, 0.0), ("Vz", 0.0, 0.0)],
                  "u": [("Ux", 90.0, 0.0), ("Uz", 0.0, 0.0)]}
# SH (src/swpc_sh/m_wav.f90:564-573)
COMPONENTS_SH = {"v": [("Vy", 90.0, 90.0)], "u": [("Uy", 90.0, 90.0)]}


def _components(dim):
    return {"3d": COMPONENTS_3D, "psv": COMPONENTS_PSV, "sh": COMPONENTS_SH}[dim]


def write_wav(odir, meta: WavMeta, stations, vel=None, disp=None, dim="3d", stlonlat=None):
    """Export receiver traces as SAC files (``wav__write`` with
    wav_format = 'sac').  ``vel``/``disp`` have shape (ntw, nst, ncomp) as
    returned by the solvers' ``receivers()``; ``stations`` are objects with
    ``x, y, z, name`` (y ignored for P-SV).  Returns the written paths."""
    comps = _components(dim)
    wdir = os.path.join(odir, "wav")
    os.makedirs(wdir, exist_ok=True)
    out = []
    for kind, arr, idep in (("v", vel, IDEP_VEL), ("u", disp, IDEP_DISP)):
        if arr is None:
            continue
        arr = np.asarray(arr, F32)
        if arr.shape[:2] != (meta.ntw, len(stations)):
            raise ValueError(f"trace array shape {arr.shape} does not match ntw={meta.ntw}, nst={len(stations)}")
        for s, st in enumerate(stations):
            lo, la = (stlonlat[s] if stlonlat is not None else (UNDEF_F, UNDEF_F))
            h0 = station_header(meta, st.name, st.x, getattr(st, "y", meta.sy0) if dim == "3d" else meta.sy0,
                                st.z, lo, la, dim)
            for c, (name, inc, az) in enumerate(comps[kind]):
                h = h0.copy()
                h.kcmpnm = name
                h.cmpinc = inc
                h.cmpaz = 0.0 if az is None else az + meta.phi
                h.idep = idep
                p = os.path.join(wdir, f"{meta.title.strip()}.{dim}.{st.name.strip()}.{name}.sac")
                write_sac(p, h, arr[:, s, c])
                out.append(p)
    return out


def wav_traces(meta: WavMeta, stations, vel=None, disp=None, dim="3d", stlonlat=None):
    """(header, samples) of every exported trace, in the reference's order:
    velocity components of all stations, then displacement (``sh_vel(j,i)``
    with the component fastest, m_wav.f90:680-700, 710-716)."""
    comps = _components(dim)
    out = []
    for kind, arr, idep in (("v", vel, IDEP_VEL), ("u", disp, IDEP_DISP)):
        if arr is None:
            continue
        arr = np.asarray(arr, F32)
        if arr.shape[:2] != (meta.ntw, len(stations)):
            raise ValueError(f"trace array shape {arr.shape} does not match ntw={meta.ntw}, nst={len(stations)}")
        for s, st in enumerate(stations):
            lo, la = (stlonlat[s] if stlonlat is not None else (UNDEF_F, UNDEF_F))
            h0 = station_header(meta, st.name, st.x, getattr(st, "y", meta.sy0) if dim == "3d" else meta.sy0,
                                st.z, lo, la, dim)
            for c, (name, inc, az) in enumerate(comps[kind]):
                h = h0.copy()
                h.kcmpnm = name
                h.cmpinc = inc
                h.cmpaz = 0.0 if az is None else az + meta.phi
                h.idep = idep
                out.append((kind, st.name.strip(), name, h, arr[:, s, c]))
    return out


def write_csf(odir, meta: WavMeta, stations, vel=None, disp=None, dim="3d", stlonlat=None, rank=0):
    """wav_format = 'csf' (export_wav__csf, m_wav.f90:798-811, and wcsf_s,
    src/shared/m_sac.f90:585-647): one file per product,
    ``<odir>/wav/<title>__<rank:05d>__.csf`` holding 'CSFD', int32 ntrace,
    int32 npts, then header + samples per trace (component fastest).  Both
    products name the same file; the reference then asks on standard input
    whether to overwrite it.  Here the first product written (velocity when
    it is on) is kept and a second one is refused with a warning, the
    reference's answer 'n'."""
    import warnings
    wdir = os.path.join(odir, "wav")
    os.makedirs(wdir, exist_ok=True)
    path = os.path.join(wdir, f"{meta.title.strip()}__{rank:05d}__.csf")
    done = False
    for arr in (vel, disp):
        if arr is None:
            continue
        if done:
            warnings.warn(f"wcsf: file {path} exists; not overwritten")
            continue
        done = True
        tr = wav_traces(meta, stations, arr if arr is vel else None, arr if arr is disp else None, dim, stlonlat)
        with open(path, "wb") as fo:
            fo.write(b"CSFD" + np.array([len(tr), meta.ntw], np.int32).tobytes())
            for _, _, _, h, d in tr:
                fo.write(h.pack())
                fo.write(np.asarray(d, F32)[:meta.ntw].tobytes())
    return path


def write_tar(odir, meta: WavMeta, stations, vel=None, disp=None, dim="3d", stlonlat=None, per="station", rank=0):
    """wav_format = 'tar_st' (one archive per station,
    ``<title>.<dim>.<stnm>.sac.tar``) or 'tar_node' (one archive per rank,
    ``<title>.<dim>.<rank:06d>.sac.tar``) of SAC members named
    ``<title>.<stnm>.<dim>.<cmp>.sac`` (m_wav.f90:717-786).  The archives
    are USTAR with mode 0644, owner 0 and mtime = exedate."""
    import io
    import tarfile
    wdir = os.path.join(odir, "wav")
    os.makedirs(wdir, exist_ok=True)
    tr = wav_traces(meta, stations, vel, disp, dim, stlonlat)
    title = meta.title.strip()
    groups = {}
    for kind, stnm, cmp, h, d in tr:
        key = stnm if per == "station" else f"{rank:06d}"
        groups.setdefault(key, []).append((stnm, cmp, h, d))
    order = [st.name.strip() for st in stations] if per == "station" else [f"{rank:06d}"]
    paths = []
    for key in order:
        if key not in groups:
            continue
        path = os.path.join(wdir, f"{title}.{dim}.{key}.sac.tar")
        with tarfile.open(path, "w", format=tarfile.USTAR_FORMAT) as tf:
            for stnm, cmp, h, d in groups[key]:
                body = h.pack() + np.asarray(d, F32)[:int(h.npts)].tobytes()
                ti = tarfile.TarInfo(f"{title}.{stnm}.{dim}.{cmp}.sac")
                ti.size, ti.mode, ti.uid, ti.gid = len(body), 0o644, 0, 0
                ti.mtime = int(meta.exedate or 0)
                tf.addfile(ti, io.BytesIO(body))
        paths.append(path)
    return paths


if sys.byteorder != "little":  # pragma: no cover
    raise ImportError("SAC export assumes a little-endian host")
