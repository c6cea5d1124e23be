"""Snapshot host logic and the oracle's snapshot products, on the CPU.

* section geometry of snap__setup (src/swpc_3d/m_snap.f90:127-162) and the
  cell order of the wbuf_* loops;
* the native header layout of write_snp_header (m_snap.f90:838-884) and a
  write/read round trip;
* the oracle's 'ps'/'v'/'u' products pinned by what the mathematics fixes:
  one-sided differences are exact on linear velocity fields, so div and
  each rotation component take known values (a swapped index or sign fails
  one of them); lam = 0 / mu = 0 cells are masked to zero; 'u' after n
  calls on a constant field is n V UC M0 dt.
"""
import numpy as np
import pytest

from openswpc_amd import problem, snapshot
from oracle.oracle import Oracle2D, Oracle3D

F32 = np.float32


@pytest.fixture(scope="module")
def case3():
    return problem.case3d(40, 36, 32, 0.5, 0.04, 4, na=6, src_depth=4.0, nst_side=2, zbeg=-2.0)


@pytest.fixture(scope="module")
def case2():
    return problem.case2d(60, 40, 0.5, 0.04, 4, na=6, src_depth=4.0, zbeg=-2.0)


def test_section_geometry(case3):
    g = case3.grid
    sg = snapshot.snap_grid3(g, idec=3, jdec=2, kdec=4)
    assert (sg.nxs, sg.nys, sg.nzs) == ((40 + 1) // 3, (36 + 1) // 2, (32 + 2) // 4)
    # coordinates of cell centres ii = i*idec - idec/2
    assert abs(sg.xsnp[0] - (g.xbeg + 1.5 * g.dx)) < 1e-6
    n1, n2, ijk = snapshot.cells3(sg, "xz")
    assert (n1, n2) == (sg.nxs, sg.nzs) and len(ijk) == n1 * n2
    # first index fastest: point 1 is (ii=2) then (ii=5) ...
    assert tuple(ijk[0]) == (3 - 1, sg.j0_xz, 4 - 2) and tuple(ijk[1]) == (6 - 1, sg.j0_xz, 4 - 2)
    assert tuple(ijk[n1]) == (2, sg.j0_xz, 8 - 2)
    n1, n2, ijk = snapshot.cells3(sg, "fs", case3.medium.kfs, case3.medium.kob)
    i, j, k = ijk[7]
    assert k == case3.medium.kfs[j + 2, i + 2] + 1
    n1, n2, ijk = snapshot.cells3(sg, "yz")
    assert np.all(ijk[:, 0] == sg.i0_yz) and (n1, n2) == (sg.nys, sg.nzs)
    # default sections: x0 = 0, y0 = 0, z0 = 10 km clipped to the model
    assert sg.i0_yz == int(np.ceil((0 - g.xbeg) / g.dx - 1e-9))
    assert sg.k0_xy == int(np.ceil((min(10.0, g.zend) - g.zbeg) / g.dz - 1e-9))


def test_header_layout_and_roundtrip(tmp_path, case3):
    g = case3.grid
    sg = snapshot.snap_grid3(g, 2, 2, 2)
    hb = snapshot.header_bytes("SWPC_3D ", "run1", 1234567, "xy", "v3", sg.xsnp, sg.ysnp, g.dt, 10, 3, 3, 6, 3,
                               139.5, 35.25, 30.0, 1.0)
    assert len(hb) == 8 + 8 + 4 + 80 + 4 + 2 + 2 + 8 + 20 + 16 + 24
    assert hb[:16] == b"STREAMIOSWPC_3D "
    assert np.frombuffer(hb[16:20], np.int32)[0] == 6
    assert hb[20:24] == b"run1" and hb[24:100] == b" " * 76
    assert np.frombuffer(hb[132:136], F32)[0] == F32(F32(g.dt) * F32(10))
    rng = np.random.default_rng(1)
    med = [rng.standard_normal((sg.nxs, sg.nys)).astype(F32) for _ in range(4)]
    f = snapshot.SnapFile(tmp_path / "t.snp", hb, med)
    frames = [[rng.standard_normal((sg.nxs, sg.nys)).astype(F32) for _ in range(3)] for _ in range(5)]
    for fr in frames:
        f.append(fr)
    f.close()
    h, m, fr = snapshot.read_snp(tmp_path / "t.snp")
    assert h["title"] == "run1" and h["exedate"] == 1234567 and h["coordinate"] == "xy" and h["datatype"] == "v3"
    assert (h["ns1"], h["ns2"], h["nmed"], h["nsnp"]) == (sg.nxs, sg.nys, 6, 3)
    assert h["clon"] == 139.5 and h["phi"] == 30.0 and h["tail"] == [1.0, 1.0, 1.0]
    assert all(np.array_equal(a, b) for a, b in zip(m, med))
    assert fr.shape == (5, 3, sg.nxs, sg.nys) and np.array_equal(fr[3, 1], frames[3][1])


def _linear_state3(ref, g, coef):
    """V fields linear in the cell indices: Vx = a i + b j + c k, etc."""
    J, I, K = np.meshgrid(np.arange(-2, g.ny + 4), np.arange(-2, g.nx + 4), np.arange(-2, g.nz + 4), indexing="ij")
    for name, (a, b, c) in zip(("Vx", "Vy", "Vz"), coef):
        ref.set_field(name, (a * I + b * J + c * K).astype(np.float64))


@pytest.mark.parametrize("bits", [32, 64])
def test_oracle_ps_exact_on_linear_fields(case3, bits):
    g = case3.grid
    ref = Oracle3D(case3, bits)
    coef = [(0.3, -0.2, 0.5), (0.7, 0.11, -0.4), (-0.6, 0.25, 0.9)]
    _linear_state3(ref, g, coef)
    (ax, bx, cx), (ay, by, cy), (az, bz, cz) = coef
    dx, dy, dz = g.dx, g.dy, g.dz
    div = ax / dx + by / dy + cz / dz
    rot_x = bz / dy - cy / dz
    rot_y = cx / dz - az / dx
    rot_z = ay / dx - bx / dy
    m = case3.medium
    # solid interior cells and air cells
    kfs = int(m.kfs[10, 10])
    pts = np.array([[10, 12, kfs + 5], [20, 8, kfs + 10], [7, 25, g.nz - 3], [15, 15, 1]], np.int32)
    sid = ref.add_snapshot(0, pts, 1, snapshot.UC3, F32(1e18))
    ref.snapshots(1)
    out = ref.snaps[sid]["frames"][0]
    scale = float(F32(snapshot.UC3) * F32(1e18) * F32(1e-3))
    want = np.array([div, rot_x, rot_y, rot_z]) * scale
    tol = 2e-6 if bits == 32 else 2e-7
    for p in range(3):
        np.testing.assert_allclose(out[:, p], want, rtol=tol * 10, atol=tol * np.abs(want).max())
    # k = 1 lies in the air above the free surface: lam = mu = 0 there
    assert m.lam[12 + 2, 15 + 2, 1 + 2] == 0 and np.all(out[:, 3] == 0)


def test_oracle_v_and_u(case3):
    g = case3.grid
    ref = Oracle3D(case3, 32)
    J, I, K = np.meshgrid(np.arange(-2, g.ny + 4), np.arange(-2, g.nx + 4), np.arange(-2, g.nz + 4), indexing="ij")
    vals = (1e-3 * (1 + I % 5), -2e-3 * (1 + J % 3), 5e-4 * (1 + K % 7))
    for name, v in zip(("Vx", "Vy", "Vz"), vals):
        ref.set_field(name, v.astype(np.float64))
    pts = np.array([[5, 6, 7], [11, 3, 20], [30, 30, 2]], np.int32)
    m0 = F32(3e16)
    sv = ref.add_snapshot(1, pts, 3, snapshot.UC3, m0)
    su = ref.add_snapshot(2, pts, 3, snapshot.UC3, m0)
    for it in range(1, 8):
        ref.snapshots(it)
    v = ref.snaps[sv]["frames"]
    u = ref.snaps[su]["frames"]
    assert len(v) == 3 and len(u) == 3          # it = 1, 4, 7
    ii, jj, kk = pts[:, 0] + 2, pts[:, 1] + 2, pts[:, 2] + 2
    V = np.stack([np.asarray(x)[jj, ii, kk] for x in vals])
    np.testing.assert_allclose(v[2], V * float(snapshot.UC3) * float(m0), rtol=1e-6)
    # u after it calls: it * V * UC * M0 * dt
    for f, it in zip(range(3), (1, 4, 7)):
        np.testing.assert_allclose(u[f], it * V * float(snapshot.UC3) * float(m0) * float(F32(g.dt)), rtol=1e-5)


def test_oracle_psv_ps(case2):
    g = case2.grid
    ref = Oracle2D(case2, 64)
    # shape (nx+6, nz+6) indexed [i+2, k+2]
    I, K = np.meshgrid(np.arange(-2, g.nx + 4), np.arange(-2, g.nz + 4), indexing="ij")
    ax, cx, az, cz = 0.4, -0.3, 0.2, 0.8
    ref.set_field("Vx", ax * I + cx * K)
    ref.set_field("Vz", az * I + cz * K)
    kfs = int(case2.medium.kfs[30 + 2])
    pts = np.array([[30, kfs + 6], [20, kfs + 12], [25, 1]], np.int32)
    sid = ref.add_snapshot(0, pts, 1, snapshot.UC2, F32(1e15))
    ref.snapshots(1)
    out = ref.snaps[sid]["frames"][0]
    scale = float(F32(snapshot.UC2) * F32(1e15) * F32(1e-3))
    want = np.array([ax / g.dx + cz / g.dz, cx / g.dz - az / g.dx]) * scale
    for p in range(2):
        np.testing.assert_allclose(out[:, p], want, rtol=1e-6)
    assert np.all(out[:, 2] == 0)


@pytest.fixture(scope="module")
def case_sh():
    return problem.case_sh(60, 40, 0.5, 0.04, 30, na=6, src_depth=4.0, nst=4, st_spacing=3.0, zbeg=-2.0)


def test_sh_header_layout(tmp_path, case_sh):
    """write_snp_header of swpc_sh (src/swpc_sh/m_snap.f90:188-232): code
    'SWPC_SH ', coordinate 'xz', snaptype 'vy', nmed 3, nsnp 1 and the
    three trailing words -12345.0."""
    g = case_sh.grid
    sg = snapshot.snap_grid2(g, 2, 3)
    hb = snapshot.header_bytes("SWPC_SH ", "shrun", 7, "xz", "vy", sg.xsnp, sg.zsnp, g.dt, 5, g.na // 2, g.na // 3,
                               3, 1, 0.0, 0.0, 0.0, -12345.0)
    f = snapshot.SnapFile(tmp_path / "s.snp", hb, [np.ones((sg.nxs, sg.nzs), F32)] * 3)
    f.append([np.full((sg.nxs, sg.nzs), 2, F32)])
    f.close()
    h, m, fr = snapshot.read_snp(tmp_path / "s.snp")
    assert h["codetype"] == "SWPC_SH "
    assert (h["coordinate"], h["datatype"], h["nmed"], h["nsnp"]) == ("xz", "vy", 3, 1)
    assert (h["ns1"], h["ns2"]) == (sg.nxs, sg.nzs) and h["tail"] == [-12345.0] * 3
    assert fr.shape == (1, 1, sg.nxs, sg.nzs) and np.all(fr == 2)


@pytest.mark.parametrize("bits", [32, 64])
def test_oracle_sh_snapshots_against_receivers(case_sh, bits):
    """The 'v' and 'u' snapshot products against the receiver traces,
    which are written by another routine (wav__store, m_wav.f90): at a
    station cell Vy x UC x M0 is the trace times UC M0, and the 'u' sum in
    real(SP) follows the receiver's FP64 displacement sum x UC M0 to FP32
    accumulation error.  ntdec_s = 3, so 'u' advances on the steps between
    frames too."""
    from oracle.oracle import OracleSH
    c = case_sh
    g = c.grid
    ref = OracleSH(c, bits)
    m0 = F32(c.glut.M0)
    pts = np.asarray(c.stations, np.int32).reshape(-1, 2)
    sv = ref.add_snapshot(1, pts, 3, snapshot.UC2, m0)
    su = ref.add_snapshot(2, pts, 3, snapshot.UC2, m0)
    ref.run(1, c.nt)
    s = float(snapshot.UC2) * float(m0)
    v, u = ref.snaps[sv]["frames"], ref.snaps[su]["frames"]
    assert len(v) == len(u) == (c.nt - 1) // 3 + 1
    assert np.abs(ref.vel).max() > 0
    for f in range(len(v)):
        it = 3 * f + 1
        np.testing.assert_allclose(v[f][0], ref.vel[it - 1] * s, rtol=1e-6, atol=0)
        np.testing.assert_allclose(u[f][0], ref.disp[it - 1] * s, rtol=2e-5, atol=1e-6 * np.abs(u[-1]).max())
