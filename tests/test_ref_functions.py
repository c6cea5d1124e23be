"""Host restatements of src/shared/m_fdtool.f90 against the reference's own
compiled code: golden vectors in tests/golden/ref_fdtool.json (written by
tests/golden/make_ref_vectors.py from oracle/_ref/libswpcref.so) and, where
the reference library is built, live calls into it.
"""
import ctypes as C
import json
import math
from pathlib import Path

import numpy as np
import pytest

from openswpc_amd import fdtool
from oracle.oracle import ref_lib

G = json.loads((Path(__file__).parent / "golden" / "ref_fdtool.json").read_text())
KINDS = ["boxcar", "triangle", "herrmann", "kupper", "cosine", "texp"]


def test_sdr2moment_golden():
    for (s, d, r), ref in zip(G["sdr2moment"]["in"], G["sdr2moment"]["out"]):
        got = fdtool.sdr2moment(s, d, r)
        assert np.allclose(got, ref, rtol=2e-6, atol=2e-7), (s, d, r)


def test_sdr2moment_invariants():
    """Double couple: zero trace and sum M_ij^2 / 2 = 1 (Aki & Richards Box 4.4);
    pure strike slip on a vertical N-S plane is mxy = 1."""
    rng = np.random.default_rng(3)
    for _ in range(50):
        mxx, myy, mzz, myz, mxz, mxy = (float(v) for v in fdtool.sdr2moment(*rng.uniform([-360, -90, -180], [360, 90, 180])))
        assert abs(mxx + myy + mzz) < 1e-6
        assert abs((mxx ** 2 + myy ** 2 + mzz ** 2) / 2 + myz ** 2 + mxz ** 2 + mxy ** 2 - 1) < 1e-5
    m = fdtool.sdr2moment(0, 90, 0)
    assert np.allclose(m, [0, 0, 0, 0, 0, 1], atol=1e-7)


@pytest.mark.parametrize("kind", range(6))
def test_momentrate_golden(kind):
    ts, tr = G["momentrate"]["ts"], G["momentrate"]["tr"]
    for t, ref in zip(G["momentrate"]["t"], G["momentrate"]["out"][str(kind)]):
        got = float(fdtool.momentrate(t, KINDS[kind], ts, tr))
        assert math.isclose(got, ref, rel_tol=3e-6, abs_tol=3e-6), (KINDS[kind], t, got, ref)


@pytest.mark.parametrize("kind", ["boxcar", "triangle", "herrmann", "kupper", "cosine"])
def test_momentrate_unit_area(kind):
    """Every moment-rate function integrates to one (m_fdtool.f90 comments)."""
    tt = np.linspace(-1, 4, 200001)
    f = np.array([float(fdtool.momentrate(t, kind, 0.2, 1.3)) for t in tt[::20]])
    assert abs(np.trapezoid(f, tt[::20]) - 1) < 2e-3


def test_ikupper_golden_and_limit():
    ts, tr = G["ikupper"]["ts"], G["ikupper"]["tr"]
    for t, ref in zip(G["ikupper"]["t"], G["ikupper"]["out"]):
        assert math.isclose(float(fdtool.ikupper(t, ts, tr)), ref, rel_tol=3e-6, abs_tol=3e-7)
    # d/dt ikupper = kupper
    t = 0.9
    e = 1e-3
    d = (float(fdtool.ikupper(t + e, 0.3, 1.7)) - float(fdtool.ikupper(t - e, 0.3, 1.7))) / (2 * e)
    assert math.isclose(d, float(fdtool.kupper(t, 0.3, 1.7)), rel_tol=1e-3)


def test_seismic_moment_golden():
    for mw, ref in zip(G["seismic_moment"]["mw"], G["seismic_moment"]["out"]):
        assert math.isclose(float(fdtool.seismic_moment(mw)), ref, rel_tol=1e-6)


def test_visco_golden():
    for v in G["visco"]:
        ts = fdtool.visco_set_relaxtime(v["nm"], v["fmin"], v["fmax"])
        assert np.allclose(ts, v["ts"], rtol=1e-6)
        z = fdtool.visco_constq_zeta(v["nm"], v["fmin"], v["fmax"], np.array(v["ts"], np.float32))
        assert math.isclose(float(z), v["zeta"], rel_tol=2e-5)
        for tau, f, ref in v["chi"]:
            got = float(fdtool.visco_chi(v["nm"], np.array(v["ts"], np.float32), tau, f))
            assert math.isclose(got, ref, rel_tol=2e-6), (v["nm"], tau, f)


def test_constq_flat_in_band():
    """With tau = nm zeta / Q (m_medium.f90:202-203) the tau-method gives
    Q^-1(w) = (tau/nm) sum_l w ts_l / (1 + w^2 ts_l^2) = zeta F(w) / Q: the
    product Q Q^-1 = zeta F(w) is close to 1 across [fmin, fmax]."""
    ts = fdtool.visco_set_relaxtime(3, 0.05, 5.0).astype(float)
    zeta = float(fdtool.visco_constq_zeta(3, 0.05, 5.0, ts))
    w = 2 * np.pi * np.logspace(np.log10(0.05), np.log10(5.0), 50)
    F = sum(w * t / (1 + (w * t) ** 2) for t in ts)
    qinv = zeta * F
    assert np.ptp(qinv) / qinv.mean() < 0.2
    assert abs(qinv.mean() - 1.0) < 0.1


def test_maps_golden():
    for m in G["maps"]:
        for x, i in zip(m["x"], m["i"]):
            assert fdtool.x2i(x, m["xbeg"], m["dx"]) == i, (x, m["xbeg"], m["dx"])
        for i, x in zip(range(-2, 12), m["i2x"]):
            assert float(fdtool.i2x(i, m["xbeg"], m["dx"])) == pytest.approx(x, rel=1e-7, abs=1e-7)


def test_vcheck_golden():
    for v in G["vcheck"]:
        vp, vs, rho, xi = v["in"]
        a, b, c, fl = fdtool.vcheck(np.array([vp]), np.array([vs]), np.array([rho]), np.array([xi]),
                                    v["vmin"], v["vmax"], v["rhomin"])
        assert np.allclose([a[0], b[0], c[0]], v["out"], rtol=1e-6)
        assert [int(x) for x in fl] == v["flags"]


def test_live_reference_library():
    """Spot-check against the amdflang build of the reference, when present."""
    L = ref_lib()
    if L is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    L.ref_momentrate.restype = C.c_float
    L.ref_visco_constq_zeta.restype = C.c_float
    rng = np.random.default_rng(9)
    for _ in range(20):
        s, d, r = rng.uniform([-360, -90, -180], [360, 90, 180])
        m = (C.c_float * 6)()
        L.ref_sdr2moment(C.c_float(s), C.c_float(d), C.c_float(r), m)
        assert np.allclose(fdtool.sdr2moment(s, d, r), list(m), rtol=2e-6, atol=2e-7)
        t = rng.uniform(-0.2, 2.5)
        assert math.isclose(float(fdtool.kupper(t, 0.1, 2.0)), L.ref_momentrate(C.c_float(t), 3, C.c_float(0.1), C.c_float(2.0)),
                            rel_tol=3e-6, abs_tol=3e-7)
