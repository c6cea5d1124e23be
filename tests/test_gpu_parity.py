"""HIP path vs the CPU oracle, through the C ABI, on seeded inputs.

Bars: FP32 kernels within relative L2 1e-5 of the FP32 oracle on
receiver traces and wavefields (the tolerance stated for the plain FP32
kernels); FP64 kernels within 1e-11 of the FP64 oracle.
"""
import numpy as np
import pytest

from openswpc_amd import problem
from openswpc_amd.solver import Swpc3D, SwpcPSV, FIELDS3, FIELDS2
from oracle.oracle import Oracle2D, Oracle3D

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
TOL64 = 1e-11
# Memory variables are small differences of O(1) terms (c1 R - c2 f dt);
# FP32 rounding-order differences show up there at ~1e-5 relative while
# the stresses and traces they feed stay within TOL32.
TOL32_R = 1e-4


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / n) if n > 0 else float(np.linalg.norm(a))


def say(msg):
    print(msg, flush=True)


CASES3 = {
    # name: (nx, ny, nz, dx, dt, nt, na, src_depth, zbeg, rmed)
    "layered": (64, 56, 48, 0.5, 0.04, 80, 8, 5.0, -2.0, None),
    "ragged": (71, 45, 39, 0.5, 0.04, 60, 6, 4.0, -1.5, None),
    "random": (80, 72, 50, 0.4, 0.03, 70, 10, 5.0, -2.0, (0.05, 2.0, 0.5)),
}


@pytest.mark.parametrize("name", list(CASES3))
@pytest.mark.parametrize("bits", [32, 64])
def test_3d_traces_and_fields(name, bits):
    nx, ny, nz, dx, dt, nt, na, sd, zb, rm = CASES3[name]
    case = problem.case3d(nx, ny, nz, dx, dt, nt, na=na, src_depth=sd, nst_side=3, zbeg=zb, rmed=rm,
                          rmed_tile=(32, 32, 32))
    say(f"3-D {name} fp{bits}: {nx}x{ny}x{nz} na={na} nt={nt}")
    ref = Oracle3D(case, bits)
    ref.run(1, nt)
    with Swpc3D(case.grid, case.medium, case.pml, case.glut, case.stations, 1, nt,
                precision="fp32" if bits == 32 else "fp64") as sim:
        sim.run(1, nt)
        vel, disp = sim.receivers()
        tol = TOL32 if bits == 32 else TOL64
        assert np.abs(ref.vel).max() > 0
        assert rel(vel, ref.vel) <= tol
        assert rel(disp, ref.disp) <= tol
        for f in ("Vx", "Vy", "Vz", "Sxx", "Syz", "Sxy", "Rxx1", "Rxz3", "axVx", "azSzz"):
            e = rel(sim.field(f), ref.field(f))
            assert e <= (TOL32_R if bits == 32 and f[0] == "R" else tol), (f, e)
        vm = sim.kernel__vmax()
        np.testing.assert_allclose(vm, ref.vmax(), rtol=1e-4 if bits == 32 else 1e-7, atol=1e-30)


@pytest.mark.parametrize("bits", [32, 64])
def test_3d_one_update_random_state(bits):
    """Every term of both half-steps: random V, S, R and ADE states, one
    stress and one velocity update, all fields compared."""
    case = problem.case3d(50, 46, 40, 0.5, 0.04, 4, na=7, src_depth=4.0, nst_side=2, zbeg=-2.0)
    say(f"3-D random-state fp{bits}")
    rng = np.random.default_rng(7)
    ref = Oracle3D(case, bits)
    with Swpc3D(case.grid, case.medium, case.pml, case.glut, case.stations, 1, 4,
                precision="fp32" if bits == 32 else "fp64") as sim:
        for f in FIELDS3:
            a = rng.standard_normal(case.grid.shape) * (1e-3 if f[0] in "RA" or f.startswith("a") else 1.0)
            # ghost cells of the memory box stay zero, as in the reference
            a[:2], a[-3:], a[:, :2], a[:, -3:], a[:, :, :2], a[:, :, -3:] = 0, 0, 0, 0, 0, 0
            a = a.astype(np.float32 if bits == 32 else np.float64)
            sim.set_field(f, a)
            ref.set_field(f, a)
        sim.kernel__update_stress()
        ref.update_stress()
        sim.source__stressglut(1)
        ref.stressglut(1)
        sim.kernel__update_vel()
        ref.update_vel()
        tol = 1e-6 if bits == 32 else 1e-13
        for f in FIELDS3:
            e = rel(sim.field(f), ref.field(f))
            assert e <= tol, (f, e)


@pytest.mark.parametrize("prec", ["fp32", "fp64", "fp32_r16"])
@pytest.mark.parametrize("dims", [(37, 29, 23), (131, 67, 45)])
def test_3d_field_roundtrip_every_field(prec, dims):
    """Every field set to distinct random values over the whole memory box
    (ghost cells included), then all read back: each equals what was set,
    so no field overlaps another and no cell lands in another's place.
    The memory variables and the stresses are stored in tiles of 256 cells
    (18 and 6 components per tile); these grids end in a partial tile and
    start their rows at different slots.  Values stored in 16 bit
    (fp32_r16 memory variables) must round-trip once rounded: the second
    read equals the first exactly, and the first is within FP16 rounding
    of the value set (inside the FP16 range of the scale)."""
    nx, ny, nz = dims
    case = problem.case3d(nx, ny, nz, 0.5, 0.04, 4, na=6, src_depth=4.0, nst_side=2, zbeg=-2.0)
    say(f"3-D field round trip {prec} {dims}")
    rng = np.random.default_rng(11)
    dt = np.float64 if prec == "fp64" else np.float32
    with Swpc3D(case.grid, case.medium, case.pml, case.glut, case.stations, 1, 4, precision=prec) as sim:
        want = {}
        for f in FIELDS3:
            if prec == "fp32_r16" and f[0] == "R":
                # inside the FP16 range of this source's scale (65504 * 2^-26,
                # about 9.8e-4); beyond it the stores saturate by design
                a = rng.uniform(-4e-4, 4e-4, case.grid.shape).astype(dt)
            else:
                a = rng.standard_normal(case.grid.shape).astype(dt)
            want[f] = a
            sim.set_field(f, a)
        for f in FIELDS3:
            got = sim.field(f)
            if prec == "fp32_r16" and f[0] == "R":
                w = want[f].astype(np.float64)
                assert np.abs(got - w).max() <= 1e-3 * np.abs(w).max(), f
                sim.set_field(f, got)
                assert np.array_equal(sim.field(f), got), f
            else:
                assert np.array_equal(got, want[f]), f


CASES2 = {
    "layered": (200, 120, 0.25, 0.02, 120, 12, 6.0, -2.0, None),
    "ragged": (131, 77, 0.25, 0.02, 90, 9, 4.0, -1.0, None),
    "random": (180, 100, 0.25, 0.02, 100, 10, 5.0, -2.0, (0.05, 1.0, 0.5)),
}


@pytest.mark.parametrize("name", list(CASES2))
@pytest.mark.parametrize("bits", [32, 64])
def test_psv_traces_and_fields(name, bits):
    nx, nz, dx, dt, nt, na, sd, zb, rm = CASES2[name]
    case = problem.case2d(nx, nz, dx, dt, nt, na=na, src_depth=sd, nst=8, st_spacing=2.0, zbeg=zb, rmed=rm)
    say(f"P-SV {name} fp{bits}: {nx}x{nz} na={na} nt={nt}")
    ref = Oracle2D(case, bits)
    ref.run(1, nt)
    with SwpcPSV(case.grid, case.medium, case.pml, case.glut, case.stations, 1, nt,
                 precision="fp32" if bits == 32 else "fp64") as sim:
        sim.run(1, nt)
        vel, disp = sim.receivers()
        tol = TOL32 if bits == 32 else TOL64
        assert np.abs(ref.vel).max() > 0
        assert rel(vel, ref.vel) <= tol
        assert rel(disp, ref.disp) <= tol
        for f in FIELDS2:
            e = rel(sim.field(f), ref.field(f))
            assert e <= (TOL32_R if bits == 32 and f[0] == "R" else tol), (f, e)
        np.testing.assert_allclose(sim.kernel__vmax(), ref.vmax(), rtol=1e-4 if bits == 32 else 1e-7)


@pytest.mark.parametrize("bits", [32, 64])
def test_psv_one_update_random_state(bits):
    case = problem.case2d(90, 60, 0.25, 0.02, 4, na=8, src_depth=3.0, nst=4, zbeg=-1.0)
    say(f"P-SV random-state fp{bits}")
    rng = np.random.default_rng(11)
    ref = Oracle2D(case, bits)
    with SwpcPSV(case.grid, case.medium, case.pml, case.glut, case.stations, 1, 4,
                 precision="fp32" if bits == 32 else "fp64") as sim:
        for f in FIELDS2:
            a = rng.standard_normal(case.grid.shape)
            a[:2], a[-3:], a[:, :2], a[:, -3:] = 0, 0, 0, 0
            a = a.astype(np.float32 if bits == 32 else np.float64)
            sim.set_field(f, a)
            ref.set_field(f, a)
        sim.kernel__update_stress()
        ref.update_stress()
        sim.source__stressglut(1)
        ref.stressglut(1)
        sim.kernel__update_vel()
        ref.update_vel()
        tol = 1e-6 if bits == 32 else 1e-13
        for f in FIELDS2:
            e = rel(sim.field(f), ref.field(f))
            assert e <= tol, (f, e)


def test_determinism():
    """Two runs of the same case are bitwise equal (race-free kernels)."""
    case = problem.case3d(48, 48, 40, 0.5, 0.04, 30, na=6, src_depth=4.0, nst_side=2, zbeg=-2.0)
    out = []
    for _ in range(2):
        with Swpc3D(case.grid, case.medium, case.pml, case.glut, case.stations, 1, 30) as sim:
            sim.run(1, 30)
            out.append(sim.field("Vz"))
    assert np.array_equal(out[0], out[1])


def test_3d_undamped_ade_skip_is_exact():
    """The PML kernels leave ADE arrays of undamped axes untouched while they
    are known to be zero; forcing the full update (by setting an ADE array
    from the host) must give bit-identical results."""
    nt = 60
    case = problem.case3d(64, 56, 48, 0.5, 0.04, nt, na=8, src_depth=5.0, nst_side=3, zbeg=-2.0)
    say("3-D ADE skip vs full PML update")
    out = []
    for full in (False, True):
        for prec in ("fp32", "fp64", "fp32_s16"):
            with Swpc3D(case.grid, case.medium, case.pml, case.glut, case.stations, 1, nt, precision=prec) as sim:
                if full:
                    sim.set_field("axVx", np.zeros(case.grid.shape))
                sim.run(1, nt)
                out.append((sim.receivers()[0], sim.field("Vx"), sim.field("aySxy")))
    for q in range(3):
        a, b = out[q], out[q + 3]
        assert np.abs(a[0]).max() > 0
        for x, y in zip(a, b):
            assert np.array_equal(x, y), q


@pytest.mark.parametrize("bits", [32, 64])
def test_3d_bodyforce(bits):
    """bf_mode: a single force injected into V after the velocity update
    (m_source.f90:835-876), traces and fields against the oracle."""
    nt = 60
    case = problem.case3d(64, 56, 48, 0.5, 0.04, nt, na=8, src_depth=5.0, nst_side=3, zbeg=-2.0, bodyforce=True)
    say(f"3-D body force fp{bits}")
    ref = Oracle3D(case, bits)
    ref.run(1, nt)
    with Swpc3D(case.grid, case.medium, case.pml, None, case.stations, 1, nt,
                precision="fp32" if bits == 32 else "fp64", bodyforce=case.bforce) as sim:
        sim.run(1, nt)
        vel, disp = sim.receivers()
        tol = TOL32 if bits == 32 else TOL64
        assert np.abs(ref.vel).max() > 0
        assert rel(vel, ref.vel) <= tol
        for f in ("Vx", "Vz", "Sxy"):
            assert rel(sim.field(f), ref.field(f)) <= tol, f


@pytest.mark.parametrize("bits", [32, 64])
def test_psv_bodyforce(bits):
    nt = 80
    case = problem.case2d(160, 80, 0.25, 0.02, nt, na=10, src_depth=5.0, zbeg=-2.0, bodyforce=True)
    say(f"P-SV body force fp{bits}")
    ref = Oracle2D(case, bits)
    ref.run(1, nt)
    with SwpcPSV(case.grid, case.medium, case.pml, None, case.stations, 1, nt,
                 precision="fp32" if bits == 32 else "fp64", bodyforce=case.bforce) as sim:
        sim.run(1, nt)
        vel, _ = sim.receivers()
        tol = TOL32 if bits == 32 else TOL64
        assert np.abs(ref.vel).max() > 0
        assert rel(vel, ref.vel) <= tol
        assert rel(sim.field("Vz"), ref.field("Vz")) <= tol


def test_divergence_detected():
    """A time step far above the stability limit: run_monitored stops with
    NumericalDivergence at a report step (m_report.f90:141-150) instead of
    running on with overflowing fields; a stable run passes its checks."""
    from openswpc_amd.solver import NumericalDivergence
    say("divergence check")
    case = problem.case3d(64, 56, 48, 0.5, 0.08, 400, na=8, src_depth=5.0, nst_side=2, zbeg=-2.0)
    assert problem.stability3(case.grid, case.medium.vmax) > 1.0
    with Swpc3D(case.grid, case.medium, case.pml, case.glut, case.stations, 1, 400) as sim:
        with pytest.raises(NumericalDivergence):
            sim.run_monitored(1, 400, ntdec_r=20)
    ok = problem.case3d(64, 56, 48, 0.5, 0.02, 100, na=8, src_depth=5.0, nst_side=2, zbeg=-2.0)
    lines = []
    with Swpc3D(ok.grid, ok.medium, ok.pml, ok.glut, ok.stations, 1, 100) as sim:
        sim.run_monitored(1, 100, ntdec_r=25, log=lines.append)
    assert len(lines) == 4


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("where", ["interior", "absorber", "edge", "bodyforce"])
def test_psv_fused_step_equals_split(bits, where):
    """swpc2d_run through the two fused step kernels against the separate
    per-routine kernels, bit for bit (traces, displacement, every field):
    a stress glut in the interior, one inside the absorber (applied by the
    absorber threads), one on the first column (its footprint leaves the
    updated cells, so the separate glut kernel applies it), and a body
    force."""
    nt = 80
    g0 = problem.case2d(160, 80, 0.25, 0.02, nt, na=10, src_depth=5.0, zbeg=-2.0).grid
    src_x = {"interior": 0.0, "absorber": g0.xbeg + 4.5 * g0.dx, "edge": g0.xbeg + 0.5 * g0.dx,
             "bodyforce": 0.0}[where]
    case = problem.case2d(160, 80, 0.25, 0.02, nt, na=10, src_depth=5.0, zbeg=-2.0, nst=8, st_spacing=2.0,
                          src_x=src_x, bodyforce=where == "bodyforce")
    src = case.bforce if where == "bodyforce" else case.glut
    say(f"P-SV fused vs split fp{bits}: {where}, source cell {tuple(src.idx[0])}")
    out = []
    for fused in (True, False):
        with SwpcPSV(case.grid, case.medium, case.pml, case.glut, case.stations, 1, nt,
                     precision="fp32" if bits == 32 else "fp64", bodyforce=case.bforce) as sim:
            sim.set_fusion(fused)
            sim.run(1, nt)
            vel, disp = sim.receivers()
            out.append((vel, disp, {f: sim.field(f) for f in FIELDS2}))
    assert np.abs(out[0][0]).max() > 0
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    for f in FIELDS2:
        assert np.array_equal(out[0][2][f], out[1][2][f]), f


def test_3d_lds_staged_velocity_is_exact():
    """The z-marching velocity update with its y-stencil rows staged in LDS
    (tuning entry 21; slower, kept as a measured alternative) gives the
    same bits as the default form on a ragged grid, for tile rows 2, 4
    and 8, after stress and velocity updates over 40 steps."""
    nt = 40
    case = problem.case3d(71, 45, 39, 0.5, 0.04, nt, na=6, src_depth=4.0, nst_side=3, zbeg=-1.5,
                          rmed=(0.05, 2.0, 0.5), rmed_tile=(32, 32, 32))
    say("3-D LDS-staged velocity vs default: 71x45x39")
    # swpc3d_set_tuning entries 0..21 (256 x 2 staged stress, velocity tile rows as below)
    base = [4, 3, 0, 4, 5, 0, 32, 0, 1, 2, 0, 1, 4, 0, 8, 0, 16, 0, 0, 1, 0, 0]
    out = []
    for vby, lds in ((4, 0), (4, 1), (2, 1), (8, 1)):
        v = list(base)
        v[12], v[21] = vby, lds
        with Swpc3D(case.grid, case.medium, case.pml, case.glut, case.stations, 1, nt) as sim:
            sim.set_tuning(*v)
            sim.run(1, nt)
            out.append((sim.receivers()[0], sim.field("Vx"), sim.field("Vy"), sim.field("Vz")))
    assert np.abs(out[0][0]).max() > 0
    for o in out[1:]:
        for x, y in zip(out[0], o):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("prec", ["fp32", "fp32_r16"])
@pytest.mark.parametrize("dims", [(71, 45, 39, 6), (300, 40, 36, 8)])
def test_3d_x_neighbours_by_shuffles_is_exact(prec, dims):
    """The 4-wide stress and velocity kernels with the x neighbours of a
    thread's cells taken from the adjacent lanes by wavefront shuffles
    (tuning entry 22, both kernels) give the same bits as with loads, for every field
    over 40 steps: interior rows whose first and last threads hold cells
    of the absorber (na not a multiple of 4) and x tiles of one and of two
    waves."""
    nx, ny, nz, na = dims
    nt = 40
    case = problem.case3d(nx, ny, nz, 0.5, 0.04, nt, na=na, src_depth=4.0, nst_side=3, zbeg=-1.5,
                          rmed=(0.05, 2.0, 0.5), rmed_tile=(32, 32, 32), st_spacing=3.0)
    say(f"3-D x neighbours by shuffles vs loads, {prec}: {nx}x{ny}x{nz} na={na}")
    # swpc3d_set_tuning entries 0..22: one-pass stress (staged), velocity tile rows 4, chunk 8
    base = [4, 3, 0, 4, 5, 0, 32, 0, 1, 2, 0, 1, 4, 0, 8, 0, 16, 0, 0, 1, 0, 0, 0]
    out = []
    for xs in (0, 3):
        v = list(base)
        v[22] = xs
        with Swpc3D(case.grid, case.medium, case.pml, case.glut, case.stations, 1, nt, precision=prec) as sim:
            sim.set_tuning(*v)
            sim.run(1, nt)
            out.append((sim.receivers()[0], {f: sim.field(f) for f in FIELDS3}))
    assert np.abs(out[0][1]["Vx"]).max() > 0
    assert np.array_equal(out[0][0], out[1][0])
    for f in FIELDS3:
        assert np.array_equal(out[0][1][f], out[1][1][f]), f
