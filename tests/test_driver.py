"""input.inf-driven host (openswpc_amd.driver) on the CPU.

* The reference's example run (example/input.inf, lhm.dat, source.dat,
  stloc.xy, held under tests/golden/example) sets up to the report header
  the reference printed for it (example/example.out:5-14): minimum and
  maximum velocity after the physical-dispersion scaling and the surface
  detection, the stability number c, the wavelength number r and the
  maximum frequency.  These depend on nz, dz, dt and the 1-D model only, so
  the check runs on a narrow 48 x 48 column of the same model.
* Parameter handling: benchmark mode, the 'uni' model with a random
  medium read from a classic netCDF volume laid out as gen_rmed3d writes
  it (m_rdrmed.f90), the source formats, 'll' stations, sdep_fit, snapshot
  switches, unsupported options.
"""
import os
import re
import shutil

import numpy as np
import pytest

from openswpc_amd import cdf, driver, fdtool, geomap, medium as md
from openswpc_amd.inifile import InputFile

F32 = np.float32
HERE = os.path.dirname(__file__)
EX = os.path.join(HERE, "golden", "example")


def example_text(**over):
    txt = open(os.path.join(EX, "input.inf"), errors="replace").read()
    for k, v in over.items():
        txt = re.sub(rf"(?m)^(\s*{re.escape(k)}\s*=\s*)\S+", lambda m: m.group(1) + str(v), txt)
    return txt


@pytest.fixture
def exdir(tmp_path, monkeypatch):
    d = tmp_path / "run"
    (d / "example").mkdir(parents=True)
    for f in ("lhm.dat", "source.dat", "stloc.xy", "stloc.ll"):
        shutil.copy(os.path.join(EX, f), d / "example" / f)
    monkeypatch.chdir(d)
    return d


def reference_report():
    out = {}
    for line in open(os.path.join(EX, "example_report.txt")):
        m = re.match(r"\s*(Stability|Wavelength|Minimum velocity|Maximum velocity|Maximum frequency)\D*:\s*([\d.]+)",
                     line)
        if m:
            out[m.group(1)] = float(m.group(2))
    return out


def test_example_report_header(exdir):
    S = driver.setup(InputFile(text=example_text(nx=48, ny=48, xbeg=-12.0, ybeg=-12.0)))
    want = reference_report()
    assert round(S.report["vmin"], 3) == want["Minimum velocity"]
    assert round(S.report["vmax"], 3) == want["Maximum velocity"]
    assert round(S.report["c"], 3) == want["Stability"]
    assert round(S.report["r"], 3) == want["Wavelength"]
    assert round(S.report["fmax"], 3) == want["Maximum frequency"]
    # the example's explosion at 2 km depth, three ocean-bottom stations
    assert len(S.stations) == 3 and [s.name for s in S.stations] == ["st01", "st02", "st03"]
    assert S.glut.M0 == pytest.approx(1e15, rel=1e-6)
    assert S.snap_cfg["fmt"] == "netcdf" and ("ob", "v") in S.snaps and ("xz", "ps") in S.snaps
    assert S.ntdec_w == 5 and S.ntdec_r == 50


def test_benchmark_mode_fixed_values(tmp_path):
    S = driver.setup(InputFile(text="benchmark_mode = .true.\nnx = 48\nny = 44\nnz = 40\nnt = 10\n"))
    g = S.grid
    assert (g.dx, g.dz, g.na) == (0.5, 0.5, 20) and F32(g.dt) == F32(0.04)
    assert g.xbeg == -12.0 and g.ybeg == -11.0 and g.zbeg == -15.0
    assert np.all(S.medium.taup == 0)          # zeta = 0: no attenuation
    assert S.glut.M0 == pytest.approx(1e15)
    m = S.glut.m[0]
    assert m[0] == pytest.approx(1 / np.sqrt(3)) and m[3] == 0


def _write_rmed3(path, vol):
    """A volume in gen_rmed3d's layout: dimensions x, y, z, coordinate
    variables, then the field (Fortran order (x, y, z))."""
    nx, ny, nz = vol.shape
    w = cdf.Writer(str(path))
    for n, s in zip("xyz", vol.shape):
        w.dim(n, s)
    for n in "xyz":
        w.var(n, "f4", [n])
    w.var("rmed", "f4", ["z", "y", "x"])
    w.enddef()
    for n, s in zip("xyz", vol.shape):
        w.put(n, np.arange(s, dtype=F32))
    w.put("rmed", np.ascontiguousarray(vol.transpose(2, 1, 0)))
    w.close()


def test_uni_rmed_psmeca_ll_stations(tmp_path):
    rng = np.random.default_rng(4)
    vol = (0.05 * rng.standard_normal((16, 12, 10))).astype(F32)
    (tmp_path / "rm").mkdir()
    _write_rmed3(tmp_path / "rm" / "a.nc", vol)
    np.testing.assert_array_equal(driver.read_rmed(str(tmp_path / "rm" / "a.nc"), 3), vol)
    src = tmp_path / "src.dat"
    # psmeca: lon lat depth mrr mtt mff mrt mrf mtf iexp
    src.write_text("# psmeca\n139.80 35.70 6.0 1.0 -0.5 -0.5 0.2 0.1 -0.3 23\n")
    st = tmp_path / "st.ll"
    st.write_text("139.7604 35.7182 0.0 A01 fsb\n139.78 35.73 1.5 A02 dep\n150.0 35.0 0.0 FAR fsb\n")
    txt = f"""
title = 'uni'
nx = 48
ny = 48
nz = 40
nt = 20
dx = 0.5
dy = 0.5
dz = 0.5
dt = 0.02
na = 8
zbeg = -2.0
vmodel_type = 'uni_rmed'
vp0 = 6.0
vs0 = 3.5
rho0 = 2.7
qp0 = 400
qs0 = 200
topo0 = 1.0
dir_rmed = '{tmp_path}/rm'
fn_rmed0 = 'a.nc'
stf_format = 'psmeca'
fn_stf = '{src}'
sw_wav_v = .true.
st_format = 'll'
fn_stloc = '{st}'
fs_v%sw = .true.
ob_u%sw = .true.
"""
    S = driver.setup(InputFile(text=txt))
    g = S.grid
    # medium: solid below topo0 = 1 km carries the tiled perturbation
    xi = md.rmed_tile(vol, -2, g.nx + 3, -2, g.ny + 3, -2, g.nz + 3)
    k = int(np.argmax(g.zc > 2.0))
    want_rho = (F32(1) + F32(0.8) * xi[:, :, k]) * F32(2.7)
    rho = S.medium.rho[:, :, k]
    inner = (slice(g.na + 3, g.ny - g.na + 2), slice(g.na + 3, g.nx - g.na + 2))
    np.testing.assert_array_equal(rho[inner], want_rho[inner])
    # ocean between 0 and 1 km, air above
    kk = int(np.argmax(g.zc > 0.3))
    assert g.zc[kk] < 1.0 and S.medium.mu[10, 10, kk] == 0 and S.medium.rho[10, 10, kk] == 1.0
    # psmeca: unit-norm tensor, moment in N m, rise time from the moment
    m = S.glut.m[0]
    assert np.isclose(np.sqrt(m[0] ** 2 + m[1] ** 2 + m[2] ** 2 + 2 * (m[3] ** 2 + m[4] ** 2 + m[5] ** 2)), np.sqrt(2),
                      rtol=1e-6)
    norm = np.sqrt(1 + 0.25 + 0.25 + 2 * (0.04 + 0.01 + 0.09)) / np.sqrt(2)
    assert S.glut.M0 == pytest.approx(norm * 1e23 * 1e-7, rel=1e-6)
    # the signs of mrf and mtf flip (myz, mxy)
    assert m[3] < 0 and m[5] > 0
    x, y = geomap.g2c(139.80, 35.70, 139.7604, 35.7182, 0.0)
    assert S.meta["sx0"] == pytest.approx(float(x)) and S.meta["evdp"] == 6.0
    # ll stations: the far one is outside the model
    assert [s.name for s in S.stations] == ["A01", "A02"]
    assert S.stlonlat[0] == (pytest.approx(139.7604, abs=1e-4), pytest.approx(35.7182, abs=1e-4))
    assert S.snaps == [("fs", "v"), ("ob", "u")]


def test_source_formats_agree(tmp_path):
    """The same double couple given as xym0dc, xymwdc, llm0dc and xym0ij
    gives the same moment tensor and moment; dsdc takes 1e9 mu D S."""
    mw = 5.0
    mo = float(fdtool.seismic_moment(mw))
    m6 = fdtool.sdr2moment(F32(30.0) - F32(10.0), 60.0, 90.0)
    lo, la = geomap.c2g(3.0, -2.0, 139.0, 36.0, 10.0)
    files = {
        "xym0dc": f"3.0 -2.0 8.0 0.5 1.2 {mo!r} 30 60 90",
        "xymwdc": f"3.0 -2.0 8.0 0.5 1.2 {mw} 30 60 90",
        "llm0dc": f"{float(lo)!r} {float(la)!r} 8.0 0.5 1.2 {mo!r} 30 60 90",
        "xym0ij": f"3.0 -2.0 8.0 0.5 1.2 {mo!r} " + " ".join(repr(float(c)) for c in m6),
    }
    from openswpc_amd.source import read_source_file
    for fmt, line in files.items():
        p = tmp_path / f"{fmt}.dat"
        p.write_text(line + "\n")
        (s,), info = read_source_file(str(p), fmt, 10.0, 139.0, 36.0, with_info=True)
        assert s.mo == pytest.approx(mo, rel=1e-6), fmt
        np.testing.assert_allclose(s.m6, m6, rtol=1e-5, atol=1e-6, err_msg=fmt)
        assert abs(s.x - 3.0) < 2e-3 and abs(s.y + 2.0) < 2e-3, fmt
        assert info["otim"] == pytest.approx(0.5)
    p = tmp_path / "ds.dat"
    p.write_text("3.0 -2.0 8.0 0.5 1.2 2.0 4.0 30 60 90\n")
    (s,) = read_source_file(str(p), "xydsdc", 10.0, 139.0, 36.0, mu_at=lambda x, y, z: 30.0)
    assert s.mo == pytest.approx(1e9 * 30.0 * 2.0 * 4.0, rel=1e-6)


def test_sdep_fit_and_unsupported(tmp_path):
    src = tmp_path / "s.dat"
    src.write_text("0 0 5.0 0.1 1.0 1e15 1 1 1 0 0 0\n")
    base = f"nx = 40\nny = 40\nnz = 40\nnt = 5\ndx = 0.5\ndy = 0.5\ndz = 0.5\ndt = 0.02\nna = 8\n" \
           f"zbeg = -2.0\nvmodel_type = 'uni'\ntopo0 = 1.5\nfn_stf = '{src}'\n"
    S = driver.setup(InputFile(text=base + "sdep_fit = 'bd0'\n"))
    k = fdtool.x2i(1.5, S.grid.zbeg, S.grid.dz)
    assert S.glut.idx[0][2] == k and S.bddep0 == 1.5
    for extra in ("pw_mode = .true.", "green_mode = .true.", "stabilize_pml = .true.", "vmodel_type = 'grd'",
                  "sw_wav_stress = .true.", "abc_type = 'cerjan'", "snp_format = 'hdf5'"):
        txt = extra + "\n" + base
        with pytest.raises(driver.InputError):
            driver.setup(InputFile(text=txt))
    with pytest.raises(driver.InputError):
        driver.setup(InputFile(text=base.replace("dt = 0.02", "dt = 0.2")))   # c > 1
