"""Snapshot streams (snap__write, src/swpc_3d/m_snap.f90; P-SV
src/swpc_psv/m_snap.f90; SH src/swpc_sh/m_snap.f90) on the GPU against the oracle, frame by frame,
through the C ABI, and the native files written from them.

Bar: relative L2 per frame <= 1e-5 (FP32 kernels) and <= 1e-6 (FP64
kernels: the products are rounded to FP32 as in the reference).
"""
import numpy as np
import pytest

from openswpc_amd import problem, snapshot
from openswpc_amd.solver import Swpc3D, SwpcPSV
from oracle.oracle import Oracle2D, Oracle3D

pytestmark = pytest.mark.gpu
F32 = np.float32


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / n) if n > 0 else float(np.linalg.norm(a))


@pytest.mark.parametrize("bits", [32, 64])
def test_3d_snapshot_frames(bits):
    nt, ntdec = 40, 5
    case = problem.case3d(64, 56, 48, 0.5, 0.04, nt, na=8, src_depth=5.0, nst_side=2, zbeg=-2.0)
    g = case.grid
    sg = snapshot.snap_grid3(g, idec=2, jdec=3, kdec=2, z0_xy=3.0, x0_yz=1.0, y0_xz=-2.0)
    print(f"3-D snapshots fp{bits}: 64x56x48, nt={nt}, every {ntdec} steps", flush=True)
    ref = Oracle3D(case, bits)
    m0 = F32(case.glut.M0)
    specs = []
    for plane in snapshot.PLANES3:
        n1, n2, ijk = snapshot.cells3(sg, plane, case.medium.kfs, case.medium.kob)
        for kind in ("ps", "v", "u"):
            specs.append((plane, kind, ijk, ref.add_snapshot(snapshot.KIND[kind], ijk, ntdec, snapshot.UC3, m0)))
    ref.run(1, nt)
    with Swpc3D(g, case.medium, case.pml, case.glut, case.stations, 1, nt,
                precision="fp32" if bits == 32 else "fp64") as sim:
        ids = [sim.add_snapshot(snapshot.KIND[k], ijk, ntdec, 16, snapshot.UC3, m0, F32(g.dt))
               for _, k, ijk, _ in specs]
        sim.run(1, nt)
        tol = 1e-5 if bits == 32 else 1e-6
        for (plane, kind, ijk, rid), sid in zip(specs, ids):
            frames = ref.snaps[rid]["frames"]
            assert len(frames) == nt // ntdec
            nvar = 4 if kind == "ps" else 3
            assert np.abs(frames[-1]).max() > 0, (plane, kind)
            for f, want in enumerate(frames):
                got = sim.get_snapshot(sid, f, nvar, len(ijk))
                e = rel(got, want)
                assert e <= tol, (plane, kind, f, e)


def test_3d_snapshot_files(tmp_path):
    """SnapshotSet: ring of 3 frames drained in chunks; files read back
    equal to the oracle frames."""
    nt, ntdec = 36, 4
    case = problem.case3d(48, 44, 40, 0.5, 0.04, nt, na=6, src_depth=4.0, nst_side=2, zbeg=-2.0)
    g = case.grid
    sg = snapshot.snap_grid3(g, 2, 2, 2)
    print("3-D snapshot files: 48x44x40", flush=True)
    ref = Oracle3D(case, 32)
    m0 = F32(case.glut.M0)
    want = {}
    for plane, kind in (("xy", "ps"), ("fs", "v"), ("xz", "u")):
        _, _, ijk = snapshot.cells3(sg, plane, case.medium.kfs, case.medium.kob)
        want[(plane, kind)] = ref.add_snapshot(snapshot.KIND[kind], ijk, ntdec, snapshot.UC3, m0)
    ref.run(1, nt)
    with Swpc3D(g, case.medium, case.pml, case.glut, case.stations, 1, nt, precision="fp32") as sim:
        ss = snapshot.SnapshotSet(sim, str(tmp_path), "t", ntdec_s=ntdec, nbuf=3, exedate=7)
        paths = {pk: ss.add3(sg, *pk, case.medium, topo=-0.0) for pk in want}
        ss.run(1, nt)
        ss.close()
    for pk, path in paths.items():
        h, med, fr = snapshot.read_snp(path)
        assert fr.shape[0] == nt // ntdec and h["codetype"] == "SWPC_3D " and h["exedate"] == 7
        for f in range(fr.shape[0]):
            w = ref.snaps[want[pk]]["frames"][f]
            got = np.stack([fr[f, v].T.ravel() for v in range(fr.shape[1])])
            assert rel(got, w) <= 1e-5, (pk, f)


@pytest.mark.parametrize("bits", [32, 64])
def test_psv_snapshot_frames(bits):
    nt, ntdec = 60, 6
    case = problem.case2d(160, 80, 0.25, 0.02, nt, na=10, src_depth=5.0, zbeg=-2.0)
    g = case.grid
    sg = snapshot.snap_grid2(g, idec=3, kdec=2)
    print(f"P-SV snapshots fp{bits}: 160x80, nt={nt}", flush=True)
    ref = Oracle2D(case, bits)
    m0 = F32(case.glut.M0)
    _, _, ik = snapshot.cells2(sg)
    rids = [ref.add_snapshot(k, ik, ntdec, snapshot.UC2, m0) for k in range(3)]
    ref.run(1, nt)
    with SwpcPSV(g, case.medium, case.pml, case.glut, case.stations, 1, nt,
                 precision="fp32" if bits == 32 else "fp64") as sim:
        ids = [sim.add_snapshot(k, ik, ntdec, 4, snapshot.UC2, m0, F32(g.dt)) for k in range(3)]
        sim.run(1, nt)
        tol = 1e-5 if bits == 32 else 1e-6
        for k, (rid, sid) in enumerate(zip(rids, ids)):
            frames = ref.snaps[rid]["frames"]
            assert np.abs(frames[-1]).max() > 0
            # the ring holds the last 4 frames; older ones are refused
            with pytest.raises(Exception):
                sim.get_snapshot(sid, 0, 2, len(ik))
            for f in range(len(frames) - 4, len(frames)):
                e = rel(sim.get_snapshot(sid, f, 2, len(ik)), frames[f])
                assert e <= tol, (k, f, e)


@pytest.mark.parametrize("bits", [32, 64])
def test_sh_snapshot_frames(bits):
    """SH 'v' and 'u' streams (src/swpc_sh/m_snap.f90:415-520) against the
    oracle's frames; the ring keeps the last nbuf frames; 'ps' does not
    exist in SH and is refused."""
    from openswpc_amd.solver import SwpcSH
    from oracle.oracle import OracleSH
    nt, ntdec = 60, 6
    case = problem.case_sh(160, 80, 0.25, 0.02, nt, na=10, src_depth=5.0, zbeg=-2.0)
    g = case.grid
    sg = snapshot.snap_grid2(g, idec=3, kdec=2)
    print(f"SH snapshots fp{bits}: 160x80, nt={nt}", flush=True)
    ref = OracleSH(case, bits)
    m0 = F32(case.glut.M0)
    _, _, ik = snapshot.cells2(sg)
    rids = [ref.add_snapshot(k, ik, ntdec, snapshot.UC2, m0) for k in (1, 2)]
    ref.run(1, nt)
    with SwpcSH(g, case.medium, case.pml, case.glut, case.stations, 1, nt,
                precision="fp32" if bits == 32 else "fp64") as sim:
        with pytest.raises(Exception):
            sim.add_snapshot(0, ik, ntdec, 4, snapshot.UC2, m0, F32(g.dt))
        ids = [sim.add_snapshot(k, ik, ntdec, 4, snapshot.UC2, m0, F32(g.dt)) for k in (1, 2)]
        sim.run(1, nt)
        tol = 1e-5 if bits == 32 else 1e-6
        for k, (rid, sid) in enumerate(zip(rids, ids)):
            frames = ref.snaps[rid]["frames"]
            assert len(frames) == (nt - 1) // ntdec + 1 and np.abs(frames[-1]).max() > 0
            with pytest.raises(Exception):
                sim.get_snapshot(sid, 0, 1, len(ik))
            for f in range(len(frames) - 4, len(frames)):
                e = rel(sim.get_snapshot(sid, f, 1, len(ik)), frames[f])
                assert e <= tol, (k, f, e)


def test_3d_netcdf_snapshots_and_peaks(tmp_path):
    """netCDF mode: frames, and the max-V/H/A fields kept on the GPU at
    every step (plus the closing update), against the oracle."""
    nt, ntdec = 30, 4
    case = problem.case3d(48, 44, 40, 0.5, 0.04, nt, na=6, src_depth=4.0, nst_side=2, zbeg=-2.0)
    g = case.grid
    sg = snapshot.snap_grid3(g, 2, 2, 2)
    print("3-D netCDF snapshots with peaks: 48x44x40", flush=True)
    ref = Oracle3D(case, 32)
    m0 = F32(case.glut.M0)
    rid = {}
    for plane, kind in (("fs", "v"), ("ob", "u"), ("xz", "ps")):
        _, _, ijk = snapshot.cells3(sg, plane, case.medium.kfs, case.medium.kob)
        rid[(plane, kind)] = ref.add_snapshot(snapshot.KIND[kind], ijk, ntdec, snapshot.UC3, m0)
        if kind != "ps":
            ref.snapshot_peaks(rid[(plane, kind)])
    ref.run(1, nt)
    ref.snapshots(nt + 1, closing=True)
    with Swpc3D(g, case.medium, case.pml, case.glut, case.stations, 1, nt, precision="fp32") as sim:
        ss = snapshot.SnapshotSet(sim, str(tmp_path), "n", ntdec_s=ntdec, nbuf=3, fmt="netcdf")
        paths = {pk: ss.add3(sg, *pk, case.medium) for pk in rid}
        ss.run(1, nt)
        ss.close(nt)
    for pk, path in paths.items():
        _, v = snapshot.read_nc_snp(path)
        names = snapshot.VNAMES3[pk[1]][0]
        frames = ref.snaps[rid[pk]]["frames"]
        for f, w in enumerate(frames):
            r = (f * ntdec + 1) // ntdec
            got = np.stack([v[nm][r].T.ravel() for nm in names])
            assert rel(got, w) <= 1e-5, (pk, f)
        if pk[1] != "ps":
            pk_got = np.stack([v[nm].T.ravel() for nm in ("max-V", "max-H", "max-A")])
            want = ref.snaps[rid[pk]]["peak"]
            assert np.abs(want).max() > 0
            assert rel(pk_got, want) <= 1e-5, pk
