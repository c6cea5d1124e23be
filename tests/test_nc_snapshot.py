"""netCDF snapshot files (snp_format = 'netcdf'), on the CPU.

SnapshotSet writes through an oracle-backed stand-in for the solver; the
files are read back with scipy.io.netcdf_file, an independent reader of
the classic format.  Checked: the dimensions, variables and attributes of
write_nc_header (src/swpc_3d/m_snap.f90:623-741), record placement and
time values of wbuf_nc (942-971), the actual_range bookkeeping of
close_nc (2185-2200), the lon/lat tables of newfile_xy_nc (743-836), and
the max-V/H/A fields of output__put_maxval (2291-2344).  The maxima of the
oracle are pinned by what they must satisfy: every frame's |Z| and
horizontal modulus stay below them, the closing update at nt+1 is
included, and max-A^2 = max-V^2 + max-H^2.
"""
import numpy as np
import pytest
from scipy.io import netcdf_file

from openswpc_amd import geomap, problem, snapshot
from oracle.oracle import Oracle2D, Oracle3D

F32 = np.float32
FILL = F32(9.9692099683868690e36)


class OracleSim:
    """The solver calls SnapshotSet makes, answered by the oracle."""

    def __init__(self, ref, case):
        self.ref, self.g, self.M0 = ref, case.grid, case.glut.M0

    def add_snapshot(self, kind, ijk, ntdec, nbuf, uc, m0, dt):
        return self.ref.add_snapshot(kind, ijk, ntdec, uc, m0)

    def get_snapshot(self, sid, frame, nvar, npts):
        return self.ref.snaps[sid]["frames"][frame]

    def run(self, it, m):
        self.ref.run(it, m)

    def snapshot_peaks(self, sid):
        self.ref.snapshot_peaks(sid)

    def finish_snapshots(self, it):
        self.ref.snapshots(it, closing=True)

    def get_peaks(self, sid, npts):
        return self.ref.snaps[sid]["peak"]


@pytest.fixture(scope="module")
def case3():
    return problem.case3d(40, 36, 32, 0.5, 0.04, 14, na=6, src_depth=4.0, nst_side=2, zbeg=-2.0)


@pytest.mark.parametrize("ntdec", [3, 1])
def test_3d_netcdf_files(tmp_path, case3, ntdec):
    nt = case3.nt
    g = case3.grid
    sg = snapshot.snap_grid3(g, 2, 3, 2)
    ref = Oracle3D(case3, 32)
    sim = OracleSim(ref, case3)
    ev = {"evlo": 139.5, "evla": 35.5, "evdp": 4.0, "evx": 0.0, "evy": 0.0}
    ss = snapshot.SnapshotSet(sim, str(tmp_path), "ncrun", ntdec_s=ntdec, nbuf=4, exedate=11, clon=139.7604,
                              clat=35.7182, phi=20.0, fmt="netcdf", ev=ev)
    paths = {pk: ss.add3(sg, *pk, case3.medium, topo=0.0) for pk in
             (("fs", "v"), ("ob", "u"), ("xz", "ps"), ("yz", "v"))}
    ss.run(1, nt)
    ss.close(nt)
    nf = (nt - 1) // ntdec + 1
    for (plane, kind), path in paths.items():
        assert path.endswith(f"ncrun.3d.{plane}.{kind}.nc")
        sid = [s.sid for s in ss.streams if (s.plane, s.kind) == (plane, kind)][0]
        frames = ref.snaps[sid]["frames"]
        assert len(frames) == nf
        with netcdf_file(path, "r", mmap=False) as f:
            n1 = sg.nxs if plane != "yz" else sg.nys
            n2 = sg.nzs if plane in ("yz", "xz") else sg.nys
            d1, d2 = {"yz": ("y", "z"), "xz": ("x", "z")}.get(plane, ("x", "y"))
            assert f.dimensions[d1] == n1 and f.dimensions[d2] == n2 and f.dimensions["t"] is None
            assert f.generated_by == b"SWPC" and f.codetype == b"SWPC_3D " and int(f.hdrver) == 6
            assert f.title == b"ncrun" and int(f.exedate) == 11
            assert (int(f.ns1), int(f.ns2)) == (n1, n2)
            assert f.coordinate == plane.encode()
            assert f.datatype == (b"ps" if kind == "ps" else (kind + "3").encode())
            assert int(f.nmed) == (6 if plane in ("xy", "fs", "ob") else 3)
            assert int(f.nsnp) == (4 if kind == "ps" else 3)
            assert F32(f.dt) == F32(F32(g.dt) * F32(ntdec))
            assert F32(f.evlo) == F32(139.5) and F32(f.phi) == F32(20.0)
            names = snapshot.VNAMES3[kind][0]
            nrec = f.variables[names[0]].shape[0]
            # the frame of it0 = r*ntdec+1 sits at record it0 // ntdec
            assert nrec == ((nf - 1) * ntdec + 1) // ntdec + 1
            t = f.variables["t"][:]
            for fr in range(nf):
                it0 = fr * ntdec + 1
                r = it0 // ntdec
                assert F32(t[r]) == F32(F32(it0) * F32(g.dt))
                for v, nm in enumerate(names):
                    got = np.asarray(f.variables[nm][r], F32)          # (n2, n1), x1 fastest
                    np.testing.assert_array_equal(got.ravel(), frames[fr][v])
            if ntdec == 1:
                assert np.all(f.variables[names[0]][0] == FILL) and F32(t[0]) == FILL
            for v, nm in enumerate(names):
                allv = np.stack([fr[v] for fr in frames])
                rng = f.variables[nm].actual_range
                assert F32(rng[0]) == min(F32(0), allv.min()) and F32(rng[1]) == max(F32(0), allv.max())
            rho = np.asarray(f.variables["rho"][:], F32)
            assert rho.shape == (n2, n1) and rho.min() > 0
            if plane in ("fs", "ob", "xy"):
                lon = np.asarray(f.variables["lon"][:], F32)
                lat = np.asarray(f.variables["lat"][:], F32)
                xs, ys = sg.xsnp, sg.ysnp
                for (a, b) in ((0, 0), (n1 - 1, n2 - 1), (3, 5)):
                    lo, la = geomap.c2g(xs[a], ys[b], 139.7604, 35.7182, 20.0)
                    assert lon[b, a] == lo and lat[b, a] == la
                assert f.variables["topo"].units == b"m"
            peaks = plane in ("fs", "ob") and kind in ("v", "u")
            assert ("max-V" in f.variables) == peaks
            if peaks:
                pk = [np.asarray(f.variables[nm][:], F32).ravel() for nm in ("max-V", "max-H", "max-A")]
                np.testing.assert_array_equal(np.stack(pk), ref.snaps[sid]["peak"])
                for fr in frames:
                    assert np.all(pk[0] >= np.abs(fr[2]))
                    assert np.all(pk[1] >= np.sqrt(fr[0] * fr[0] + fr[1] * fr[1]))
                np.testing.assert_allclose(pk[2] ** 2, pk[0] ** 2 + pk[1] ** 2, rtol=1e-6)
                assert f.variables["max-V"].units == (b"m/s" if kind == "v" else b"m")


def test_closing_update_enters_the_maxima(case3):
    """The maxima include the state after the last step (the wbuf_*(nt+1)
    call of snap__closefiles), which no frame shows."""
    nt = case3.nt
    ref = Oracle3D(case3, 32)
    sg = snapshot.snap_grid3(case3.grid, 2, 2, 2)
    _, _, ijk = snapshot.cells3(sg, "fs", case3.medium.kfs, case3.medium.kob)
    sid = ref.add_snapshot(1, ijk, 1, snapshot.UC3, F32(case3.glut.M0))
    ref.snapshot_peaks(sid)
    ref.run(1, nt)
    before = ref.snaps[sid]["peak"].copy()
    ref.snapshots(nt + 1, closing=True)
    after = ref.snaps[sid]["peak"]
    assert len(ref.snaps[sid]["frames"]) == nt
    assert np.all(after >= before) and np.any(after > before)
    # with frames every step the pre-closing maxima are the frame maxima
    fr = np.stack(ref.snaps[sid]["frames"])
    np.testing.assert_array_equal(before[0], np.abs(fr[:, 2]).max(axis=0))


def test_psv_netcdf_file(tmp_path):
    nt, ntdec = 12, 4
    case = problem.case2d(60, 40, 0.5, 0.04, nt, na=6, src_depth=4.0, zbeg=-2.0)
    ref = Oracle2D(case, 32)
    sim = OracleSim(ref, case)
    sg = snapshot.snap_grid2(case.grid, 2, 2)
    ss = snapshot.SnapshotSet(sim, str(tmp_path), "p", ntdec_s=ntdec, nbuf=2, fmt="netcdf")
    path = ss.add2(sg, "v", case.medium)
    ss.run(1, nt)
    ss.close(nt)
    assert path.endswith("p.psv.xz.v.nc")
    with netcdf_file(path, "r", mmap=False) as f:
        assert f.codetype == b"SWPC_PSV" and f.datatype == b"v2" and int(f.nsnp) == 2
        assert set(f.variables) == {"x", "z", "t", "rho", "lambda", "mu", "Vx", "Vz"}
        frames = ref.snaps[ss.streams[0].sid]["frames"]
        for fr in range(len(frames)):
            r = (fr * ntdec + 1) // ntdec
            np.testing.assert_array_equal(np.asarray(f.variables["Vz"][r], F32).ravel(), frames[fr][1])
