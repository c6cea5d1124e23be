"""Host-side mirror of the reference's hot-path module interface, driving the
HIP kernels through the C ABI.

``Swpc3D`` / ``SwpcPSV`` own one device context each and expose the
reference's per-step routines under their reference names
(``kernel__update_stress``, ``absorb__update_stress`` folded in,
``source__stressglut``, ``kernel__update_vel``, ``source__bodyforce``,
``wav__store``, ``kernel__vmax``) plus ``step``/``run`` for whole steps in
the order of src/swpc_3d/main.f90:125-144 (P-SV: src/swpc_psv/main.f90:95-114).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib, medium as md
from ._lib import SwpcError, check, lib, ptr

UC3 = 1e-15   # src/swpc_3d/m_global.f90:29
UC2 = 1e-12   # src/swpc_psv/m_global.f90:29

FIELDS3 = ["Vx", "Vy", "Vz", "Sxx", "Syy", "Szz", "Syz", "Sxz", "Sxy"] + \
    [f"R{c}{m}" for c in ("xx", "yy", "zz", "yz", "xz", "xy") for m in (1, 2, 3)] + \
    ["axVx", "ayVx", "azVx", "axVy", "ayVy", "azVy", "axVz", "ayVz", "azVz",
     "axSxx", "aySxy", "azSxz", "axSxy", "aySyy", "azSyz", "axSxz", "aySyz", "azSzz"]
FIELDS2 = ["Vx", "Vz", "Sxx", "Szz", "Sxz"] + \
    [f"R{c}{m}" for c in ("xx", "zz", "xz") for m in (1, 2, 3)] + \
    ["axVx", "azVx", "axVz", "azVz", "axSxx", "azSxz", "axSxz", "azSzz"]

PRECISION = {"fp32": _lib.SWPC_FP32, "fp64": _lib.SWPC_FP64, "fp32_s16": _lib.SWPC_FP32_S16,
             "fp32_r16": _lib.SWPC_FP32_R16}


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class NumericalDivergence(SwpcError):
    """Raised by run_monitored when max|V| * UC exceeds 1e5 (the check of
    report__progress, src/swpc_3d/m_report.f90:141-150) or is not finite."""


DIVERGENCE_TOL = 1e5


def _run_monitored(sim, it0, nsteps, ntdec_r, log):
    """report__progress (m_report.f90:120-185) around the time loop: at every
    it with mod(it, ntdec_r) = 0, before that step's updates, max|V| at the
    ocean-bottom/surface layer is checked for divergence and reported as
    physical velocities (times UC*M0), with the time per step and the ETA."""
    import math
    import time
    t0 = time.perf_counter()
    it = it0
    end = it0 + nsteps
    while it < end:
        nxt = (it // ntdec_r + 1) * ntdec_r if it % ntdec_r else it
        if nxt > it:
            m = min(nxt, end) - it
            sim.run(it, m)
            it += m
            continue
        vm = sim.kernel__vmax().astype(np.float64)
        amp = float(vm.max()) * sim.UC
        if not math.isfinite(amp) or not np.all(np.isfinite(vm)) or amp > DIVERGENCE_TOL:
            raise NumericalDivergence(f"numerical divergence detected at it={it} with max amp = {amp:.3e} "
                                      "[(m/s)/moment]")
        if log is not None:
            el = time.perf_counter() - t0
            done = it - it0
            per = el / done if done else 0.0
            eta = per * (end - it)
            h, r = divmod(int(eta), 3600)
            va = vm * sim.UC * sim.M0
            log(f"  it={it:07d},{per:6.3f} s/loop, eta {h:03d}:{r // 60:02d}:{r % 60:02d}, ("
                + " ".join(f"{v:9.2E}" for v in va) + " )")
        m = min(ntdec_r, end - it)
        sim.run(it, m)
        it += m


class Swpc3D:
    """One 3-D simulation resident on the GPU."""

    def __init__(self, grid, medium, pml, glut=None, stations=None, ntdec_w=1, nt=None,
                 precision="fp32", bodyforce=None):
        self.g = grid
        self.nt = nt
        self.precision = precision
        L = lib()
        d = _lib.Desc3()
        d.nx, d.ny, d.nz, d.na, d.nm = grid.nx, grid.ny, grid.nz, grid.na, 3
        d.precision = PRECISION[precision]
        d.dx, d.dy, d.dz, d.dt = grid.dx, grid.dy, grid.dz, grid.dt
        raw = isinstance(medium, md.RawMedium3D)
        if raw:
            ts, zeta3, cc = medium.constants()
        else:
            ts = medium.ts
        for m in range(3):
            d.ts[m] = float(ts[m])
        h = C.c_void_p()
        check(L.swpc3d_create(C.byref(d), C.byref(h)), "swpc3d_create")
        self.h = h
        m = medium
        if raw:
            # medium__setup on the GPU; kfs/kob and vmin/vmax come back
            arrs = [_f32(a) for a in (m.rho, m.lam, m.mu, m.qp, m.qs)]
            kfs = np.zeros((grid.ny + 6, grid.nx + 6), np.int32)
            kob = np.zeros_like(kfs)
            mm = np.zeros(2, np.float32)
            check(L.swpc3d_setup_medium(h, *[ptr(a, C.c_float) for a in arrs], float(zeta3), float(cc.real),
                                        float(cc.imag), ptr(kfs, C.c_int), ptr(kob, C.c_int), ptr(mm, C.c_float)),
                  "setup_medium")
            m.kfs, m.kob, m.vmin, m.vmax = kfs, kob, float(mm[0]), float(mm[1])
            arrs = None
        else:
            self._keep = [_f32(a) for a in (m.bx, m.by, m.bz, m.lam, m.mu, m.taup, m.taus, m.muyz, m.muxz, m.muxy)]
            check(L.swpc3d_set_medium(h, *[ptr(a, C.c_float) for a in self._keep]), "set_medium")
            rho = _f32(m.rho)
            check(L.swpc3d_set_density(h, ptr(rho, C.c_float)), "set_density")
            s = [_i32(a) for a in (m.kfs_top, m.kfs_bot, m.kob_top, m.kob_bot, m.kob)]
            check(L.swpc3d_set_surface(h, *[ptr(a, C.c_int) for a in s]), "set_surface")
            self._keep = None
        gp = [_f32(a) for a in pml]
        check(L.swpc3d_set_pml(h, *[ptr(a, C.c_float) for a in gp]), "set_pml")
        self.M0 = 1.0
        self.UC = UC3
        if glut is not None and len(glut.idx):
            idx = _i32(glut.idx)
            i, j, k = _i32(idx[:, 0]), _i32(idx[:, 1]), _i32(idx[:, 2])
            tab = _f64(glut.table)
            check(L.swpc3d_set_sources(h, len(idx), ptr(i, C.c_int), ptr(j, C.c_int), ptr(k, C.c_int),
                                       ptr(_f64(glut.m), C.c_double), tab.shape[0], ptr(tab, C.c_double)),
                  "set_sources")
            self.M0 = glut.M0
        if bodyforce is not None and len(bodyforce.idx):
            idx = _i32(bodyforce.idx)
            i, j, k = _i32(idx[:, 0]), _i32(idx[:, 1]), _i32(idx[:, 2])
            tab = _f64(bodyforce.table)
            check(L.swpc3d_set_bodyforce(h, len(idx), ptr(i, C.c_int), ptr(j, C.c_int), ptr(k, C.c_int),
                                         ptr(_f64(bodyforce.m), C.c_double), tab.shape[0], ptr(tab, C.c_double)),
                  "set_bodyforce")
            self.M0 = bodyforce.M0
            # bf_mode unit conversion UC = UC * 10**3 in real(SP) (m_source.f90:190-195)
            self.UC = float(np.float32(np.float32(UC3) * np.float32(1000)))
        self.nst = 0
        self.ntw = 0
        self.ntdec_w = ntdec_w
        if stations is not None and len(stations):
            self.set_receivers(stations)

    def set_receivers(self, stations):
        """Receiver cells (n, 3) of (i, j, k), sampled every ntdec_w steps."""
        L = lib()
        st = _i32(stations)
        from .source import ntw_of
        self.nst = len(st)
        self.ntw = ntw_of(self.nt, self.ntdec_w)
        check(L.swpc3d_set_receivers(self.h, self.nst, ptr(_i32(st[:, 0]), C.c_int), ptr(_i32(st[:, 1]), C.c_int),
                                     ptr(_i32(st[:, 2]), C.c_int), self.ntdec_w, self.ntw), "set_receivers")

    def medium_array(self, name):
        """One medium array read back from the device in the memory-box
        (j, i, k) shape: bx by bz lam mu taup taus muyz muxz muxy or rho."""
        names = ("bx", "by", "bz", "lam", "mu", "taup", "taus", "muyz", "muxz", "muxy", "rho")
        out = np.zeros((self.g.ny + 6, self.g.nx + 6, self.g.nz + 6), np.float32)
        check(lib().swpc3d_get_medium(self.h, names.index(name), ptr(out, C.c_float)), "get_medium")
        return out

    # -- reference routine names ------------------------------------------
    def kernel__update_stress(self):
        check(lib().swpc3d_update_stress(self.h), "update_stress")

    def kernel__update_vel(self):
        check(lib().swpc3d_update_vel(self.h), "update_vel")

    def source__stressglut(self, it):
        check(lib().swpc3d_stressglut(self.h, it), "stressglut")

    def source__bodyforce(self, it):
        check(lib().swpc3d_bodyforce(self.h, it), "bodyforce")

    def wav__store(self, it):
        check(lib().swpc3d_store_receivers(self.h, it), "store_receivers")

    def kernel__vmax(self):
        out = np.zeros(3, np.float32)
        check(lib().swpc3d_vmax(self.h, ptr(out, C.c_float)), "vmax")
        return out

    def run_monitored(self, it0, nsteps, ntdec_r=100, log=None):
        """run() with the reference's progress report and divergence check
        every ntdec_r steps; raises NumericalDivergence."""
        _run_monitored(self, it0, nsteps, ntdec_r, log)

    # -- whole steps ------------------------------------------------------
    def step(self, it):
        check(lib().swpc3d_step(self.h, it), "step")

    def run(self, it0, nsteps):
        check(lib().swpc3d_run(self.h, it0, nsteps), "run")

    def sync(self):
        check(lib().swpc3d_sync(self.h), "sync")

    # -- gathers ------------------------------------------------------------
    def receivers(self, scaled=False):
        """(vel[ntw, nst, 3] float32, disp[ntw, nst, 3] float64); scaled=True
        multiplies by M0*UC*1e9 (the reference's nm/s and nm)."""
        vel = np.zeros((self.ntw, self.nst, 3), np.float32)
        disp = np.zeros((self.ntw, self.nst, 3), np.float64)
        if self.nst:
            check(lib().swpc3d_get_receivers(self.h, ptr(vel, C.c_float), ptr(disp, C.c_double)), "get_receivers")
        if scaled:
            s = self.M0 * self.UC * 1e9
            return (vel * np.float32(s)).astype(np.float32), disp * s
        return vel, disp

    def add_snapshot(self, kind, ijk, ntdec_s, nbuf, uc, m0, dt):
        """Register a snapshot stream over 1-based (i,j,k) cells (snap__write,
        m_snap.f90); returns its id."""
        ijk = _i32(np.asarray(ijk).reshape(-1, 3))
        sid = C.c_int()
        check(lib().swpc3d_add_snapshot(self.h, int(kind), len(ijk), ptr(ijk, C.c_int), int(ntdec_s), int(nbuf),
                                        float(uc), float(m0), float(dt), C.byref(sid)), "add_snapshot")
        return sid.value

    def get_snapshot(self, sid, frame, nvar, npts):
        out = np.zeros((nvar, npts), np.float32)
        check(lib().swpc3d_get_snapshot(self.h, int(sid), int(frame), ptr(out, C.c_float)), "get_snapshot")
        return out

    def snapshot_peaks(self, sid):
        """Keep max-V/H/A of a 'v' or 'u' stream from the next step on."""
        check(lib().swpc3d_snapshot_peaks(self.h, int(sid)), "snapshot_peaks")

    def finish_snapshots(self, it):
        """Closing update at step number it (snap__closefiles in netCDF mode)."""
        check(lib().swpc3d_finish_snapshots(self.h, int(it)), "finish_snapshots")

    def get_peaks(self, sid, npts):
        out = np.zeros((3, npts), np.float32)
        check(lib().swpc3d_get_peaks(self.h, int(sid), ptr(out, C.c_float)), "get_peaks")
        return out

    def field(self, name):
        out = np.zeros(self.g.shape, np.float64)
        check(lib().swpc3d_get_field(self.h, FIELDS3.index(name), ptr(out, C.c_double)), "get_field")
        return out

    def set_field(self, name, a):
        a = _f64(a)
        assert a.shape == self.g.shape
        check(lib().swpc3d_set_field(self.h, FIELDS3.index(name), ptr(a, C.c_double)), "set_field")

    def profile(self, enable=True):
        check(lib().swpc3d_profile(self.h, 1 if enable else 0), "profile")

    def profile_read(self):
        ms = np.zeros(5, np.float64)
        cnt = np.zeros(5, np.int64)
        check(lib().swpc3d_profile_read(self.h, ptr(ms, C.c_double), ptr(cnt, C.c_longlong)), "profile_read")
        return ms, cnt

    def set_tuning(self, *p):
        """Interior launch configuration, see swpc3d_set_tuning."""
        a = np.array(p, np.int32)
        check(lib().swpc3d_set_tuning(self.h, ptr(a, C.c_int), len(a)), "set_tuning")

    def memory(self):
        b = C.c_longlong()
        check(lib().swpc3d_memory(self.h, C.byref(b)), "memory")
        return b.value

    def set_learned_boundary(self, params):
        """Learned absorbing boundary in place of the PML (grid with na = 0):
        params[10][swpc_lab_nparam()] FP32 (openswpc_amd.lab), used in BF16;
        None switches it off."""
        if params is None:
            check(lib().swpc3d_set_lab(self.h, None), "set_lab")
            return
        p = _f32(params)
        check(lib().swpc3d_set_lab(self.h, ptr(p, C.c_float)), "set_lab")

    def set_learned_boundary_sponge(self, sponge):
        """Strip sponge strength of the learned boundary (0 = none, default 0.02)."""
        check(lib().swpc3d_set_lab_sponge(self.h, float(sponge)), "swpc3d_set_lab_sponge")

    def close(self):
        if getattr(self, "h", None):
            lib().swpc3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class SwpcPSV:
    """One 2-D P-SV simulation resident on the GPU."""

    def __init__(self, grid, medium, pml, glut=None, stations=None, ntdec_w=1, nt=None,
                 precision="fp32", bodyforce=None):
        self.g = grid
        self.nt = nt
        L = lib()
        d = _lib.Desc2()
        d.nx, d.nz, d.na, d.nm = grid.nx, grid.nz, grid.na, 3
        d.precision = PRECISION[precision]
        d.dx, d.dz, d.dt = grid.dx, grid.dz, grid.dt
        for m in range(3):
            d.ts[m] = float(medium.ts[m])
        h = C.c_void_p()
        check(L.swpc2d_create(C.byref(d), C.byref(h)), "swpc2d_create")
        self.h = h
        m = medium
        arrs = [_f32(a) for a in (m.rho, m.lam, m.mu, m.taup, m.taus)]
        check(L.swpc2d_set_medium(h, *[ptr(a, C.c_float) for a in arrs]), "set_medium")
        s = [_i32(a) for a in (m.kfs_top, m.kfs_bot, m.kob_top, m.kob_bot, m.kob)]
        check(L.swpc2d_set_surface(h, *[ptr(a, C.c_int) for a in s]), "set_surface")
        gp = [_f32(a) for a in pml]
        check(L.swpc2d_set_pml(h, *[ptr(a, C.c_float) for a in gp]), "set_pml")
        self.M0 = 1.0
        self.UC = UC2
        if glut is not None and len(glut.idx):
            idx = _i32(glut.idx)
            tab = _f64(glut.table)
            check(L.swpc2d_set_sources(h, len(idx), ptr(_i32(idx[:, 0]), C.c_int), ptr(_i32(idx[:, 1]), C.c_int),
                                       ptr(_f64(glut.m), C.c_double), tab.shape[0], ptr(tab, C.c_double)),
                  "set_sources")
            self.M0 = glut.M0
        if bodyforce is not None and len(bodyforce.idx):
            idx = _i32(bodyforce.idx)
            tab = _f64(bodyforce.table)
            check(L.swpc2d_set_bodyforce(h, len(idx), ptr(_i32(idx[:, 0]), C.c_int), ptr(_i32(idx[:, 1]), C.c_int),
                                         ptr(_f64(bodyforce.m), C.c_double), tab.shape[0], ptr(tab, C.c_double)),
                  "set_bodyforce")
            self.M0 = bodyforce.M0
            # bf_mode unit conversion in real(SP) (src/swpc_psv/m_source.f90:157-160)
            self.UC = float(np.float32(np.float32(UC2) * np.float32(1000)))
        self.nst = 0
        self.ntw = 0
        if stations is not None and len(stations):
            st = _i32(stations)
            from .source import ntw_of
            self.nst = len(st)
            self.ntw = ntw_of(nt, ntdec_w)
            check(L.swpc2d_set_receivers(h, self.nst, ptr(_i32(st[:, 0]), C.c_int), ptr(_i32(st[:, 1]), C.c_int),
                                         ntdec_w, self.ntw), "set_receivers")

    def kernel__update_stress(self):
        check(lib().swpc2d_update_stress(self.h), "update_stress")

    def kernel__update_vel(self):
        check(lib().swpc2d_update_vel(self.h), "update_vel")

    def source__stressglut(self, it):
        check(lib().swpc2d_stressglut(self.h, it), "stressglut")

    def source__bodyforce(self, it):
        check(lib().swpc2d_bodyforce(self.h, it), "bodyforce")

    def wav__store(self, it):
        check(lib().swpc2d_store_receivers(self.h, it), "store_receivers")

    def kernel__vmax(self):
        out = np.zeros(2, np.float32)
        check(lib().swpc2d_vmax(self.h, ptr(out, C.c_float)), "vmax")
        return out

    def run_monitored(self, it0, nsteps, ntdec_r=100, log=None):
        """run() with the reference's progress report and divergence check
        every ntdec_r steps (src/swpc_psv/m_report.f90); raises
        NumericalDivergence."""
        _run_monitored(self, it0, nsteps, ntdec_r, log)

    def step(self, it):
        check(lib().swpc2d_step(self.h, it), "step")

    def run(self, it0, nsteps):
        check(lib().swpc2d_run(self.h, it0, nsteps), "run")

    def sync(self):
        check(lib().swpc2d_sync(self.h), "sync")

    def receivers(self, scaled=False):
        vel = np.zeros((self.ntw, self.nst, 2), np.float32)
        disp = np.zeros((self.ntw, self.nst, 2), np.float64)
        if self.nst:
            check(lib().swpc2d_get_receivers(self.h, ptr(vel, C.c_float), ptr(disp, C.c_double)), "get_receivers")
        if scaled:
            s = self.M0 * self.UC * 1e9
            return (vel * np.float32(s)).astype(np.float32), disp * s
        return vel, disp

    def add_snapshot(self, kind, ik, ntdec_s, nbuf, uc, m0, dt):
        """Register a P-SV snapshot stream over 1-based (i,k) cells
        (src/swpc_psv/m_snap.f90); returns its id."""
        ik = _i32(np.asarray(ik).reshape(-1, 2))
        sid = C.c_int()
        check(lib().swpc2d_add_snapshot(self.h, int(kind), len(ik), ptr(ik, C.c_int), int(ntdec_s), int(nbuf),
                                        float(uc), float(m0), float(dt), C.byref(sid)), "add_snapshot")
        return sid.value

    def get_snapshot(self, sid, frame, nvar, npts):
        out = np.zeros((nvar, npts), np.float32)
        check(lib().swpc2d_get_snapshot(self.h, int(sid), int(frame), ptr(out, C.c_float)), "get_snapshot")
        return out

    def field(self, name):
        out = np.zeros(self.g.shape, np.float64)
        check(lib().swpc2d_get_field(self.h, FIELDS2.index(name), ptr(out, C.c_double)), "get_field")
        return out

    def set_field(self, name, a):
        a = _f64(a)
        assert a.shape == self.g.shape
        check(lib().swpc2d_set_field(self.h, FIELDS2.index(name), ptr(a, C.c_double)), "set_field")

    def set_graph(self, steps):
        """Steps per captured HIP graph in run() (0: step by step)."""
        check(lib().swpc2d_set_graph(self.h, int(steps)), "set_graph")

    def set_fusion(self, on):
        """Steps in the two fused kernels (True, the default) or in the
        separate per-routine kernels (False); the results are the same."""
        check(lib().swpc2d_set_fusion(self.h, int(bool(on))), "set_fusion")

    def close(self):
        if getattr(self, "h", None):
            lib().swpc2d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


FIELDS_SH = ["Vy", "Syz", "Sxy", "Ryz1", "Ryz2", "Ryz3", "Rxy1", "Rxy2", "Rxy3", "axVy", "azVy", "axSxy", "azSyz"]


class SwpcSH:
    """One 2-D SH simulation resident on the GPU (swpc_sh)."""

    def __init__(self, grid, medium, pml, glut=None, stations=None, ntdec_w=1, nt=None,
                 precision="fp32", bodyforce=None):
        self.g = grid
        self.nt = nt
        L = lib()
        d = _lib.Desc2()
        d.nx, d.nz, d.na, d.nm = grid.nx, grid.nz, grid.na, 3
        d.precision = PRECISION[precision]
        d.dx, d.dz, d.dt = grid.dx, grid.dz, grid.dt
        for m in range(3):
            d.ts[m] = float(medium.ts[m])
        h = C.c_void_p()
        check(L.swpcsh_create(C.byref(d), C.byref(h)), "swpcsh_create")
        self.h = h
        m = medium
        arrs = [_f32(a) for a in (m.rho, m.mu, m.taus)]
        check(L.swpcsh_set_medium(h, *[ptr(a, C.c_float) for a in arrs]), "set_medium")
        s = [_i32(a) for a in (m.kfs_top, m.kfs_bot, m.kob_top, m.kob_bot, m.kob)]
        check(L.swpcsh_set_surface(h, *[ptr(a, C.c_int) for a in s]), "set_surface")
        gp = [_f32(a) for a in pml]
        check(L.swpcsh_set_pml(h, *[ptr(a, C.c_float) for a in gp]), "set_pml")
        self.M0 = 1.0
        self.UC = UC2
        if glut is not None and len(glut.idx):
            idx = _i32(glut.idx)
            tab = _f64(glut.table)
            check(L.swpcsh_set_sources(h, len(idx), ptr(_i32(idx[:, 0]), C.c_int), ptr(_i32(idx[:, 1]), C.c_int),
                                       ptr(_f64(glut.m), C.c_double), tab.shape[0], ptr(tab, C.c_double)),
                  "set_sources")
            self.M0 = glut.M0
        if bodyforce is not None and len(bodyforce.idx):
            idx = _i32(bodyforce.idx)
            tab = _f64(bodyforce.table)
            check(L.swpcsh_set_bodyforce(h, len(idx), ptr(_i32(idx[:, 0]), C.c_int), ptr(_i32(idx[:, 1]), C.c_int),
                                         ptr(_f64(bodyforce.m), C.c_double), tab.shape[0], ptr(tab, C.c_double)),
                  "set_bodyforce")
            self.M0 = bodyforce.M0
            # bf_mode unit conversion in real(SP) (src/swpc_sh/m_source.f90:160)
            self.UC = float(np.float32(np.float32(UC2) * np.float32(1000)))
        self.nst = 0
        self.ntw = 0
        if stations is not None and len(stations):
            st = _i32(stations)
            from .source import ntw_of
            self.nst = len(st)
            self.ntw = ntw_of(nt, ntdec_w)
            check(L.swpcsh_set_receivers(h, self.nst, ptr(_i32(st[:, 0]), C.c_int), ptr(_i32(st[:, 1]), C.c_int),
                                         ntdec_w, self.ntw), "set_receivers")

    def kernel__update_stress(self):
        check(lib().swpcsh_update_stress(self.h), "update_stress")

    def kernel__update_vel(self):
        check(lib().swpcsh_update_vel(self.h), "update_vel")

    def source__stressglut(self, it):
        check(lib().swpcsh_stressglut(self.h, it), "stressglut")

    def source__bodyforce(self, it):
        check(lib().swpcsh_bodyforce(self.h, it), "bodyforce")

    def wav__store(self, it):
        check(lib().swpcsh_store_receivers(self.h, it), "store_receivers")

    def kernel__vmax(self):
        out = np.zeros(1, np.float32)
        check(lib().swpcsh_vmax(self.h, ptr(out, C.c_float)), "vmax")
        return out

    def add_snapshot(self, kind, ik, ntdec_s, nbuf, uc, m0, dt):
        """Register an SH snapshot stream ('v' or 'u') over 1-based (i,k)
        cells (src/swpc_sh/m_snap.f90); returns its id."""
        ik = _i32(np.asarray(ik).reshape(-1, 2))
        sid = C.c_int()
        check(lib().swpcsh_add_snapshot(self.h, int(kind), len(ik), ptr(ik, C.c_int), int(ntdec_s), int(nbuf),
                                        float(uc), float(m0), float(dt), C.byref(sid)), "add_snapshot")
        return sid.value

    def snapshots(self, it):
        check(lib().swpcsh_snapshots(self.h, it), "snapshots")

    def get_snapshot(self, sid, frame, nvar, npts):
        assert nvar == 1
        out = np.zeros((1, npts), np.float32)
        check(lib().swpcsh_get_snapshot(self.h, int(sid), int(frame), ptr(out, C.c_float)), "get_snapshot")
        return out

    def step(self, it):
        check(lib().swpcsh_step(self.h, it), "step")

    def run(self, it0, nsteps):
        check(lib().swpcsh_run(self.h, it0, nsteps), "run")

    def set_graph(self, steps):
        """Steps per captured HIP graph in run() (0: step by step)."""
        check(lib().swpcsh_set_graph(self.h, int(steps)), "set_graph")

    def sync(self):
        check(lib().swpcsh_sync(self.h), "sync")

    def receivers(self, scaled=False):
        """vel[itw][n] (Vy), disp[itw][n]; scaled: x M0 UC 1e9 as wav__store writes them."""
        vel = np.zeros((self.ntw, self.nst), np.float32)
        disp = np.zeros((self.ntw, self.nst), np.float64)
        if self.nst:
            check(lib().swpcsh_get_receivers(self.h, ptr(vel, C.c_float), ptr(disp, C.c_double)), "get_receivers")
        if scaled:
            s = self.M0 * self.UC * 1e9
            return (vel * np.float32(s)).astype(np.float32), disp * s
        return vel, disp

    def field(self, name):
        out = np.zeros(self.g.shape, np.float64)
        check(lib().swpcsh_get_field(self.h, FIELDS_SH.index(name), ptr(out, C.c_double)), "get_field")
        return out

    def set_field(self, name, a):
        a = _f64(a)
        assert a.shape == self.g.shape
        check(lib().swpcsh_set_field(self.h, FIELDS_SH.index(name), ptr(a, C.c_double)), "set_field")

    def close(self):
        if getattr(self, "h", None):
            lib().swpcsh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
