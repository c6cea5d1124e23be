"""ctypes harness of the CPU oracle (oracle/swpc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke()
as the checker, and by bench.py's cpu_baseline leg.  Never by the product.

``Oracle3D`` / ``Oracle2D`` own numpy arrays in the reference's Fortran
layout (memory box, k fastest) and advance them with the C restatement of
the reference loops in FP32 (liboracle32.so) or FP64 (liboracle64.so).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_libs = {}


def olib(bits=32):
    if bits not in _libs:
        p = HERE / f"liboracle{bits}.so"
        if not p.exists():
            raise RuntimeError(f"{p} missing: run `make -C oracle`")
        _libs[bits] = C.CDLL(str(p))
    return _libs[bits]


def ref_lib():
    """oracle/_ref/libswpcref.so (reference Fortran modules) or None."""
    p = HERE / "_ref" / "libswpcref.so"
    if not p.exists():
        return None
    if "ref" not in _libs:
        _libs["ref"] = C.CDLL(str(p))
    return _libs["ref"]


_P = C.c_void_p
_PF = C.POINTER(C.c_float)
_PI = C.POINTER(C.c_int)


def _o3d_struct(real):
    PR = C.POINTER(real)

    class O3D(C.Structure):
        _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("na", C.c_int), ("nm", C.c_int),
                    ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_float)] + \
                   [(n, PR) for n in ("Vx", "Vy", "Vz", "Sxx", "Syy", "Szz", "Syz", "Sxz", "Sxy")] + \
                   [(n, PR) for n in ("Rxx", "Ryy", "Rzz", "Ryz", "Rxz", "Rxy")] + \
                   [(n, _PF) for n in ("bx", "by", "bz", "lam", "mu", "taup", "taus", "muyz", "muxz", "muxy")] + \
                   [(n, _PI) for n in ("kfs_top", "kfs_bot", "kob_top", "kob_bot", "kob")] + \
                   [(n, _PF) for n in ("gxc", "gxe", "gyc", "gye", "gzc", "gze")] + \
                   [("ade", PR * 18), ("ts", _PF)]
    return O3D


def _o2d_struct(real):
    PR = C.POINTER(real)

    class O2D(C.Structure):
        _fields_ = [("nx", C.c_int), ("nz", C.c_int), ("na", C.c_int), ("nm", C.c_int),
                    ("dx", C.c_double), ("dz", C.c_double), ("dt", C.c_float)] + \
                   [(n, PR) for n in ("Vx", "Vz", "Sxx", "Szz", "Sxz")] + \
                   [(n, PR) for n in ("Rxx", "Rzz", "Rxz")] + \
                   [(n, _PF) for n in ("rho", "lam", "mu", "taup", "taus")] + \
                   [(n, _PI) for n in ("kfs_top", "kfs_bot", "kob_top", "kob_bot", "kob")] + \
                   [(n, _PF) for n in ("gxc", "gxe", "gzc", "gze")] + \
                   [("ade", PR * 8), ("ts", _PF)]
    return O2D


FIELDS3 = ["Vx", "Vy", "Vz", "Sxx", "Syy", "Szz", "Syz", "Sxz", "Sxy"]
R3 = ["Rxx", "Ryy", "Rzz", "Ryz", "Rxz", "Rxy"]
ADE3 = ["axVx", "ayVx", "azVx", "axVy", "ayVy", "azVy", "axVz", "ayVz", "azVz",
        "axSxx", "aySxy", "azSxz", "axSxy", "aySyy", "azSyz", "axSxz", "aySyz", "azSzz"]
FIELDS2 = ["Vx", "Vz", "Sxx", "Szz", "Sxz"]
R2 = ["Rxx", "Rzz", "Rxz"]
ADE2 = ["axVx", "azVx", "axVz", "azVz", "axSxx", "azSxz", "axSxz", "azSzz"]


class Oracle3D:
    def __init__(self, case, bits=32):
        self.case = case
        self.bits = bits
        self.L = olib(bits)
        real = C.c_float if bits == 32 else C.c_double
        self.dtype = np.float32 if bits == 32 else np.float64
        g, m = case.grid, case.medium
        shp = g.shape
        self.f = {n: np.zeros(shp, self.dtype) for n in FIELDS3}
        self.R = {n: np.zeros(shp + (3,), self.dtype) for n in R3}   # (m, box), m fastest
        self.A = {n: np.zeros(shp, self.dtype) for n in ADE3}
        self.med = {n: np.ascontiguousarray(getattr(m, n), np.float32)
                    for n in ("bx", "by", "bz", "lam", "mu", "taup", "taus", "muyz", "muxz", "muxy")}
        self.surf = {n: np.ascontiguousarray(getattr(m, n), np.int32)
                     for n in ("kfs_top", "kfs_bot", "kob_top", "kob_bot", "kob")}
        self.prof = [np.ascontiguousarray(a, np.float32) for a in case.pml]
        self.ts = np.ascontiguousarray(m.ts, np.float32)
        S = _o3d_struct(real)
        s = S()
        s.nx, s.ny, s.nz, s.na, s.nm = g.nx, g.ny, g.nz, g.na, 3
        s.dx, s.dy, s.dz, s.dt = g.dx, g.dy, g.dz, g.dt
        PR = C.POINTER(real)
        for n in FIELDS3:
            setattr(s, n, self.f[n].ctypes.data_as(PR))
        for n in R3:
            setattr(s, n, self.R[n].ctypes.data_as(PR))
        for n, a in self.med.items():
            setattr(s, n, a.ctypes.data_as(_PF))
        for n, a in self.surf.items():
            setattr(s, n, a.ctypes.data_as(_PI))
        for n, a in zip(("gxc", "gxe", "gyc", "gye", "gzc", "gze"), self.prof):
            setattr(s, n, a.ctypes.data_as(_PF))
        for q, n in enumerate(ADE3):
            s.ade[q] = self.A[n].ctypes.data_as(PR)
        s.ts = self.ts.ctypes.data_as(_PF)
        self.s = s
        gl = case.glut
        self.nsrc = len(gl.idx) if gl is not None else 0
        if self.nsrc:
            self.isrc = np.ascontiguousarray(gl.idx[:, 0], np.int32)
            self.jsrc = np.ascontiguousarray(gl.idx[:, 1], np.int32)
            self.ksrc = np.ascontiguousarray(gl.idx[:, 2], np.int32)
            self.m6 = np.ascontiguousarray(gl.m, np.float64)
            self.tab = np.ascontiguousarray(gl.table, np.float64)
        bf = getattr(case, "bforce", None)
        self.nbf = len(bf.idx) if bf is not None else 0
        if self.nbf:
            self.bf_idx = [np.ascontiguousarray(bf.idx[:, q], np.int32) for q in range(3)]
            self.bf_f = np.ascontiguousarray(bf.m, np.float64)
            self.bf_tab = np.ascontiguousarray(bf.table, np.float64)
        st = case.stations
        self.nst = len(st)
        if self.nst:
            self.ist = np.ascontiguousarray(st[:, 0], np.int32)
            self.jst = np.ascontiguousarray(st[:, 1], np.int32)
            self.kst = np.ascontiguousarray(st[:, 2], np.int32)
            from openswpc_amd.source import ntw_of
            self.ntw = ntw_of(case.nt, case.ntdec_w)
            self.vel = np.zeros((self.ntw, self.nst, 3), np.float32)
            self.disp = np.zeros((self.ntw, self.nst, 3), np.float64)
            self.uacc = np.zeros((self.nst, 3), np.float64)

    def _ps(self):
        return C.byref(self.s)

    def receivers(self, it):
        if not self.nst:
            return
        dec = self.case.ntdec_w
        store = (it - 1) % dec == 0 and (it - 1) // dec < self.ntw
        vel = np.zeros((self.nst, 3), np.float32)
        self.L.o3d_receivers(self._ps(), self.nst, self.ist.ctypes.data_as(_PI), self.jst.ctypes.data_as(_PI),
                             self.kst.ctypes.data_as(_PI), vel.ctypes.data_as(_PF),
                             self.uacc.ctypes.data_as(C.POINTER(C.c_double)))
        if store:
            itw = (it - 1) // dec
            self.vel[itw] = vel
            self.disp[itw] = self.uacc

    def update_stress(self):
        self.L.o3d_update_stress(self._ps())
        self.L.o3d_absorb_stress(self._ps())

    def update_vel(self):
        self.L.o3d_update_vel(self._ps())
        self.L.o3d_absorb_vel(self._ps())

    def stressglut(self, it):
        if self.nsrc:
            row = np.ascontiguousarray(self.tab[it - 1])
            PD = C.POINTER(C.c_double)
            self.L.o3d_stressglut(self._ps(), self.nsrc, self.isrc.ctypes.data_as(_PI), self.jsrc.ctypes.data_as(_PI),
                                  self.ksrc.ctypes.data_as(_PI), self.m6.ctypes.data_as(PD), row.ctypes.data_as(PD))

    def add_snapshot(self, kind, ijk, ntdec_s, uc, m0):
        """A snapshot stream over 1-based (i,j,k) cells (m_snap.f90); frames
        are kept in self.snaps[id]['frames']."""
        ijk = np.ascontiguousarray(ijk, np.int32).reshape(-1, 3)
        n = len(ijk)
        nvar = 4 if kind == 0 else 3
        if not hasattr(self, "snaps"):
            self.snaps = []
        self.snaps.append({"kind": kind, "ijk": ijk, "ntdec": ntdec_s, "uc": np.float32(uc), "m0": np.float32(m0),
                           "acc": np.zeros(3 * n, np.float32), "nvar": nvar, "frames": [], "peak": None})
        return len(self.snaps) - 1

    def snapshot_peaks(self, sid):
        """Keep the running maxima of a 'v' or 'u' stream at every call:
        max |Z|, max sqrt(X^2+Y^2) and the root of their squares, in SP
        (wbuf_fs_v and siblings, src/swpc_3d/m_snap.f90:1690-1703)."""
        sn = self.snaps[sid]
        assert sn["kind"] in (1, 2)
        sn["peak"] = np.zeros((3, len(sn["ijk"])), np.float32)

    @staticmethod
    def _peak(pk, b):
        pk[0] = np.maximum(pk[0], np.abs(b[2]))
        pk[1] = np.maximum(pk[1], np.sqrt(b[0] * b[0] + b[1] * b[1]))
        pk[2] = np.sqrt(pk[0] * pk[0] + pk[1] * pk[1])

    def snapshots(self, it, closing=False):
        """snap__write(it); closing=True is the wbuf_*(nt+1) update of
        snap__closefiles (m_snap.f90:2215-2231): sums and maxima advance,
        no frame is kept."""
        for sn in getattr(self, "snaps", []):
            due = (it - 1) % sn["ntdec"] == 0 and not closing
            if not due and sn["kind"] != 2 and sn["peak"] is None:
                continue
            if closing and sn["kind"] == 0:
                continue
            n = len(sn["ijk"])
            out = np.zeros(sn["nvar"] * n, np.float32)
            self.L.o3d_snapshot(self._ps(), sn["kind"], n, sn["ijk"].ctypes.data_as(_PI), C.c_float(sn["uc"]),
                                C.c_float(sn["m0"]), sn["acc"].ctypes.data_as(_PF), out.ctypes.data_as(_PF))
            if sn["peak"] is not None:
                self._peak(sn["peak"], out.reshape(sn["nvar"], n))
            if due:
                sn["frames"].append(out.reshape(sn["nvar"], n))

    def bodyforce(self, it):
        if self.nbf:
            row = np.ascontiguousarray(self.bf_tab[it - 1])
            PD = C.POINTER(C.c_double)
            i, j, k = self.bf_idx
            self.L.o3d_bodyforce(self._ps(), self.nbf, i.ctypes.data_as(_PI), j.ctypes.data_as(_PI),
                                 k.ctypes.data_as(_PI), self.bf_f.ctypes.data_as(PD), row.ctypes.data_as(PD))

    def step(self, it):
        """main.f90:125-144 order (receivers and snapshots sampled before
        the updates)."""
        self.receivers(it)
        self.snapshots(it)
        self.update_stress()
        self.stressglut(it)
        self.update_vel()
        self.bodyforce(it)

    def run(self, it0, n):
        for it in range(it0, it0 + n):
            self.step(it)

    def field(self, name):
        if name in self.f:
            return self.f[name]
        if name in self.A:
            return self.A[name]
        c, m = name[:3], int(name[3:])
        return self.R[c][..., m - 1]

    def set_field(self, name, a):
        if name in self.f:
            self.f[name][...] = a
        elif name in self.A:
            self.A[name][...] = a
        else:
            c, m = name[:3], int(name[3:])
            self.R[c][..., m - 1] = a

    def vmax(self):
        out = np.zeros(3, np.float32)
        self.L.o3d_vmax(self._ps(), out.ctypes.data_as(_PF))
        return out

    # -- learned absorbing boundary (o3d_lab_*) ---------------------------
    def set_lab(self, params, sponge=0.02):
        """Learned boundary on (params[10][NP] as given, already rounded as
        the time step uses them; strip sponge strength `sponge`) or off
        (None); needs na = 0."""
        if params is None:
            self.lab = None
            return
        assert self.case.grid.na == 0
        self.lab = np.ascontiguousarray(params, np.float32)
        self.lab_sponge = float(sponge)
        self.rho = np.ascontiguousarray(self.case.medium.rho, np.float32)

    def lab_fill(self, phase):
        self.L.o3d_lab_fill(self._ps(), phase, self.lab.ctypes.data_as(_PF), self.rho.ctypes.data_as(_PF),
                            C.c_float(self.lab_sponge))

    def lab_gather_begin(self, oi, oj, nxs, nys, nzs):
        """Training gather for the inner box (oi, oj, nxs, nys, nzs); the
        previous-ghost state of every net starts at zero."""
        self.rho = np.ascontiguousarray(self.case.medium.rho, np.float32)
        self.gdom = (oi, oj, nxs, nys, nzs)
        ext = [nys * nzs, nys * nzs, nxs * nzs, nxs * nzs, nxs * nys]
        self.gprev = {(f, ph): np.zeros((ext[f], 5), np.float32) for f in range(5) for ph in (0, 1)}

    def lab_gather(self, phase):
        """Records (column order) of the five faces of one phase: {face: (n, 72)}."""
        oi, oj, nxs, nys, nzs = self.gdom
        out = {}
        for f in range(5):
            pv = self.gprev[(f, phase)]
            buf = np.zeros((len(pv), 72), np.float32)
            n = self.L.o3d_lab_gather(self._ps(), phase, f, self.rho.ctypes.data_as(_PF), oi, oj, nxs, nys, nzs,
                                      pv.ctypes.data_as(_PF), buf.ctypes.data_as(_PF))
            out[f] = buf[:n]
        return out

    def lab_step(self, it):
        """step() of a learned-boundary grid: ghost fill before each
        half-step (the HIP step order)."""
        self.receivers(it)
        self.snapshots(it)
        self.lab_fill(0)
        self.update_stress()
        self.stressglut(it)
        self.lab_fill(1)
        self.update_vel()
        self.bodyforce(it)


class Oracle2D:
    def __init__(self, case, bits=32):
        self.case = case
        self.bits = bits
        self.L = olib(bits)
        real = C.c_float if bits == 32 else C.c_double
        self.dtype = np.float32 if bits == 32 else np.float64
        g, m = case.grid, case.medium
        shp = g.shape
        self.f = {n: np.zeros(shp, self.dtype) for n in FIELDS2}
        self.R = {n: np.zeros(shp + (3,), self.dtype) for n in R2}
        self.A = {n: np.zeros(shp, self.dtype) for n in ADE2}
        self.med = {n: np.ascontiguousarray(getattr(m, n), np.float32) for n in ("rho", "lam", "mu", "taup", "taus")}
        self.surf = {n: np.ascontiguousarray(getattr(m, n), np.int32)
                     for n in ("kfs_top", "kfs_bot", "kob_top", "kob_bot", "kob")}
        self.prof = [np.ascontiguousarray(a, np.float32) for a in case.pml]
        self.ts = np.ascontiguousarray(m.ts, np.float32)
        S = _o2d_struct(real)
        s = S()
        s.nx, s.nz, s.na, s.nm = g.nx, g.nz, g.na, 3
        s.dx, s.dz, s.dt = g.dx, g.dz, g.dt
        PR = C.POINTER(real)
        for n in FIELDS2:
            setattr(s, n, self.f[n].ctypes.data_as(PR))
        for n in R2:
            setattr(s, n, self.R[n].ctypes.data_as(PR))
        for n, a in self.med.items():
            setattr(s, n, a.ctypes.data_as(_PF))
        for n, a in self.surf.items():
            setattr(s, n, a.ctypes.data_as(_PI))
        for n, a in zip(("gxc", "gxe", "gzc", "gze"), self.prof):
            setattr(s, n, a.ctypes.data_as(_PF))
        for q, n in enumerate(ADE2):
            s.ade[q] = self.A[n].ctypes.data_as(PR)
        s.ts = self.ts.ctypes.data_as(_PF)
        self.s = s
        gl = case.glut
        self.nsrc = len(gl.idx) if gl is not None else 0
        if self.nsrc:
            self.isrc = np.ascontiguousarray(gl.idx[:, 0], np.int32)
            self.ksrc = np.ascontiguousarray(gl.idx[:, 1], np.int32)
            self.m3 = np.ascontiguousarray(gl.m, np.float64)
            self.tab = np.ascontiguousarray(gl.table, np.float64)
        bf = getattr(case, "bforce", None)
        self.nbf = len(bf.idx) if bf is not None else 0
        if self.nbf:
            self.bf_idx = [np.ascontiguousarray(bf.idx[:, q], np.int32) for q in range(2)]
            self.bf_f = np.ascontiguousarray(bf.m, np.float64)
            self.bf_tab = np.ascontiguousarray(bf.table, np.float64)
        st = case.stations
        self.nst = len(st)
        if self.nst:
            self.ist = np.ascontiguousarray(st[:, 0], np.int32)
            self.kst = np.ascontiguousarray(st[:, 1], np.int32)
            from openswpc_amd.source import ntw_of
            self.ntw = ntw_of(case.nt, case.ntdec_w)
            self.vel = np.zeros((self.ntw, self.nst, 2), np.float32)
            self.disp = np.zeros((self.ntw, self.nst, 2), np.float64)
            self.uacc = np.zeros((self.nst, 2), np.float64)

    def _ps(self):
        return C.byref(self.s)

    def receivers(self, it):
        if not self.nst:
            return
        dec = self.case.ntdec_w
        store = (it - 1) % dec == 0 and (it - 1) // dec < self.ntw
        vel = np.zeros((self.nst, 2), np.float32)
        self.L.o2d_receivers(self._ps(), self.nst, self.ist.ctypes.data_as(_PI), self.kst.ctypes.data_as(_PI),
                             vel.ctypes.data_as(_PF), self.uacc.ctypes.data_as(C.POINTER(C.c_double)))
        if store:
            itw = (it - 1) // dec
            self.vel[itw] = vel
            self.disp[itw] = self.uacc

    def update_stress(self):
        self.L.o2d_update_stress(self._ps())
        self.L.o2d_absorb_stress(self._ps())

    def update_vel(self):
        self.L.o2d_update_vel(self._ps())
        self.L.o2d_absorb_vel(self._ps())

    def stressglut(self, it):
        if self.nsrc:
            row = np.ascontiguousarray(self.tab[it - 1])
            PD = C.POINTER(C.c_double)
            self.L.o2d_stressglut(self._ps(), self.nsrc, self.isrc.ctypes.data_as(_PI), self.ksrc.ctypes.data_as(_PI),
                                  self.m3.ctypes.data_as(PD), row.ctypes.data_as(PD))

    def add_snapshot(self, kind, ik, ntdec_s, uc, m0):
        """A P-SV snapshot stream over 1-based (i,k) cells (m_snap.f90)."""
        ik = np.ascontiguousarray(ik, np.int32).reshape(-1, 2)
        if not hasattr(self, "snaps"):
            self.snaps = []
        self.snaps.append({"kind": kind, "ik": ik, "ntdec": ntdec_s, "uc": np.float32(uc), "m0": np.float32(m0),
                           "acc": np.zeros(2 * len(ik), np.float32), "frames": []})
        return len(self.snaps) - 1

    def snapshots(self, it):
        for sn in getattr(self, "snaps", []):
            due = (it - 1) % sn["ntdec"] == 0
            if not due and sn["kind"] != 2:
                continue
            n = len(sn["ik"])
            out = np.zeros(2 * n, np.float32)
            self.L.o2d_snapshot(self._ps(), sn["kind"], n, sn["ik"].ctypes.data_as(_PI), C.c_float(sn["uc"]),
                                C.c_float(sn["m0"]), sn["acc"].ctypes.data_as(_PF), out.ctypes.data_as(_PF))
            if due:
                sn["frames"].append(out.reshape(2, n))

    def bodyforce(self, it):
        if self.nbf:
            row = np.ascontiguousarray(self.bf_tab[it - 1])
            PD = C.POINTER(C.c_double)
            i, k = self.bf_idx
            self.L.o2d_bodyforce(self._ps(), self.nbf, i.ctypes.data_as(_PI), k.ctypes.data_as(_PI),
                                 self.bf_f.ctypes.data_as(PD), row.ctypes.data_as(PD))

    def step(self, it):
        """src/swpc_psv/main.f90:95-114 order; the body force (added after the
        velocity update there, before the absorber) commutes with the
        absorber's disjoint-cell update."""
        self.snapshots(it)
        self.receivers(it)
        self.update_stress()
        self.stressglut(it)
        self.update_vel()
        self.bodyforce(it)

    def run(self, it0, n):
        for it in range(it0, it0 + n):
            self.step(it)

    def field(self, name):
        if name in self.f:
            return self.f[name]
        if name in self.A:
            return self.A[name]
        c, m = name[:3], int(name[3:])
        return self.R[c][..., m - 1]

    def set_field(self, name, a):
        if name in self.f:
            self.f[name][...] = a
        elif name in self.A:
            self.A[name][...] = a
        else:
            c, m = name[:3], int(name[3:])
            self.R[c][..., m - 1] = a

    def vmax(self):
        out = np.zeros(2, np.float32)
        self.L.o2d_vmax(self._ps(), out.ctypes.data_as(_PF))
        return out


FIELDS_SH = ["Vy", "Syz", "Sxy"]
R_SH = ["Ryz", "Rxy"]
ADE_SH = ["axVy", "azVy", "axSxy", "azSyz"]


def _osh_struct(real):
    PR = C.POINTER(real)

    class OSH(C.Structure):
        _fields_ = [("nx", C.c_int), ("nz", C.c_int), ("na", C.c_int), ("nm", C.c_int),
                    ("dx", C.c_double), ("dz", C.c_double), ("dt", C.c_float)] + \
                   [(n, PR) for n in FIELDS_SH] + [(n, PR) for n in R_SH] + \
                   [(n, _PF) for n in ("rho", "mu", "taus")] + \
                   [(n, _PI) for n in ("kfs_top", "kfs_bot", "kob_top", "kob_bot", "kob")] + \
                   [(n, _PF) for n in ("gxc", "gxe", "gzc", "gze")] + \
                   [("ade", PR * 4), ("ts", _PF)]
    return OSH


class OracleSH:
    """The SH restatement (osh_* in swpc_oracle.c, src/swpc_sh) on a
    problem.CaseSH: fields over the (i, k) memory box (nx+6, nz+6)."""

    def __init__(self, case, bits=32):
        self.case = case
        self.bits = bits
        self.L = olib(bits)
        real = C.c_float if bits == 32 else C.c_double
        self.dtype = np.float32 if bits == 32 else np.float64
        g, m = case.grid, case.medium
        shp = g.shape
        self.f = {n: np.zeros(shp, self.dtype) for n in FIELDS_SH}
        self.R = {n: np.zeros(shp + (3,), self.dtype) for n in R_SH}
        self.A = {n: np.zeros(shp, self.dtype) for n in ADE_SH}
        self.med = {n: np.ascontiguousarray(getattr(m, n), np.float32) for n in ("rho", "mu", "taus")}
        self.surf = {n: np.ascontiguousarray(getattr(m, n), np.int32)
                     for n in ("kfs_top", "kfs_bot", "kob_top", "kob_bot", "kob")}
        self.prof = [np.ascontiguousarray(a, np.float32) for a in case.pml]
        self.ts = np.ascontiguousarray(m.ts, np.float32)
        s = _osh_struct(real)()
        s.nx, s.nz, s.na, s.nm = g.nx, g.nz, g.na, 3
        s.dx, s.dz, s.dt = g.dx, g.dz, g.dt
        PR = C.POINTER(real)
        for n in FIELDS_SH:
            setattr(s, n, self.f[n].ctypes.data_as(PR))
        for n in R_SH:
            setattr(s, n, self.R[n].ctypes.data_as(PR))
        for n, a in self.med.items():
            setattr(s, n, a.ctypes.data_as(_PF))
        for n, a in self.surf.items():
            setattr(s, n, a.ctypes.data_as(_PI))
        for n, a in zip(("gxc", "gxe", "gzc", "gze"), self.prof):
            setattr(s, n, a.ctypes.data_as(_PF))
        for q, n in enumerate(ADE_SH):
            s.ade[q] = self.A[n].ctypes.data_as(PR)
        s.ts = self.ts.ctypes.data_as(_PF)
        self.s = s
        PD = C.POINTER(C.c_double)
        self.L.osh_vmax.restype = C.c_float
        gl = case.glut
        self.nsrc = len(gl.idx) if gl is not None else 0
        if self.nsrc:
            self.src = [np.ascontiguousarray(gl.idx[:, q], np.int32) for q in range(2)]
            self.m2 = np.ascontiguousarray(gl.m, np.float64)
            self.tab = np.ascontiguousarray(gl.table, np.float64)
        bf = case.bforce
        self.nbf = len(bf.idx) if bf is not None else 0
        if self.nbf:
            self.bf_idx = [np.ascontiguousarray(bf.idx[:, q], np.int32) for q in range(2)]
            self.bf_f = np.ascontiguousarray(bf.m, np.float64)
            self.bf_tab = np.ascontiguousarray(bf.table, np.float64)
        st = case.stations
        self.nst = len(st)
        if self.nst:
            from openswpc_amd.source import ntw_of
            self.ist = np.ascontiguousarray(st[:, 0], np.int32)
            self.kst = np.ascontiguousarray(st[:, 1], np.int32)
            self.ntw = ntw_of(case.nt, case.ntdec_w)
            self.vel = np.zeros((self.ntw, self.nst), np.float32)
            self.disp = np.zeros((self.ntw, self.nst), np.float64)
            self.uacc = np.zeros(self.nst, np.float64)
        self._PD = PD

    def _ps(self):
        return C.byref(self.s)

    def receivers(self, it):
        if not self.nst:
            return
        dec = self.case.ntdec_w
        vel = np.zeros(self.nst, np.float32)
        self.L.osh_receivers(self._ps(), self.nst, self.ist.ctypes.data_as(_PI), self.kst.ctypes.data_as(_PI),
                             vel.ctypes.data_as(_PF), self.uacc.ctypes.data_as(self._PD))
        if (it - 1) % dec == 0 and (it - 1) // dec < self.ntw:
            self.vel[(it - 1) // dec] = vel
            self.disp[(it - 1) // dec] = self.uacc

    def update_stress(self):
        self.L.osh_update_stress(self._ps())
        self.L.osh_absorb_stress(self._ps())

    def update_vel(self):
        self.L.osh_update_vel(self._ps())
        self.L.osh_absorb_vel(self._ps())

    def stressglut(self, it):
        if self.nsrc:
            row = np.ascontiguousarray(self.tab[it - 1])
            i, k = self.src
            self.L.osh_stressglut(self._ps(), self.nsrc, i.ctypes.data_as(_PI), k.ctypes.data_as(_PI),
                                  self.m2.ctypes.data_as(self._PD), row.ctypes.data_as(self._PD))

    def bodyforce(self, it):
        if self.nbf:
            row = np.ascontiguousarray(self.bf_tab[it - 1])
            i, k = self.bf_idx
            self.L.osh_bodyforce(self._ps(), self.nbf, i.ctypes.data_as(_PI), k.ctypes.data_as(_PI),
                                 self.bf_f.ctypes.data_as(self._PD), row.ctypes.data_as(self._PD))

    def add_snapshot(self, kind, ik, ntdec_s, uc, m0):
        """An SH snapshot stream ('v' = 1 or 'u' = 2) over 1-based (i,k)
        cells (src/swpc_sh/m_snap.f90); frames in self.snaps[id]['frames']."""
        assert kind in (1, 2)
        ik = np.ascontiguousarray(ik, np.int32).reshape(-1, 2)
        if not hasattr(self, "snaps"):
            self.snaps = []
        self.snaps.append({"kind": kind, "ik": ik, "ntdec": ntdec_s, "uc": np.float32(uc), "m0": np.float32(m0),
                           "acc": np.zeros(len(ik), np.float32), "frames": []})
        return len(self.snaps) - 1

    def snapshots(self, it):
        for sn in getattr(self, "snaps", []):
            due = (it - 1) % sn["ntdec"] == 0
            if not due and sn["kind"] != 2:
                continue
            n = len(sn["ik"])
            out = np.zeros(n, np.float32)
            self.L.osh_snapshot(self._ps(), sn["kind"], n, sn["ik"].ctypes.data_as(_PI), C.c_float(sn["uc"]),
                                C.c_float(sn["m0"]), sn["acc"].ctypes.data_as(_PF), out.ctypes.data_as(_PF))
            if due:
                sn["frames"].append(out.reshape(1, n))

    def step(self, it):
        """src/swpc_sh/main.f90:97-114 order."""
        self.receivers(it)
        self.snapshots(it)
        self.update_stress()
        self.stressglut(it)
        self.update_vel()
        self.bodyforce(it)

    def run(self, it0, n):
        for it in range(it0, it0 + n):
            self.step(it)

    def field(self, name):
        if name in self.f:
            return self.f[name]
        if name in self.A:
            return self.A[name]
        return self.R[name[:3]][..., int(name[3:]) - 1]

    def set_field(self, name, a):
        if name in self.f:
            self.f[name][...] = a
        elif name in self.A:
            self.A[name][...] = a
        else:
            self.R[name[:3]][..., int(name[3:]) - 1] = a

    def vmax(self):
        return float(self.L.osh_vmax(self._ps()))
