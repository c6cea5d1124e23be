"""medium__setup on the GPU (swpc3d_setup_medium) against the host assembly
(medium.medium_setup3, the restatement of src/swpc_3d/m_medium.f90:36-467)
on the same velocity model: the arrays the time step reads, the surface
maps, vmin/vmax, and a short run of each."""
import numpy as np
import pytest

from openswpc_amd import medium as md, problem
from openswpc_amd._lib import SwpcError
from openswpc_amd.solver import Swpc3D

pytestmark = pytest.mark.gpu

# bit-identical to the host form; the rest go through the relaxation
# factor 1/Re(1/sqrt(1 - cc tau)), which the host forms in complex FP32
# steps and the device in FP64 rounded once (a few FP32 ulps apart)
EXACT = ("bx", "by", "bz", "taup", "taus", "rho")
RELAXED = ("lam", "mu", "muyz", "muxz", "muxy")


def _pair(na, zbeg, benchmark, seed=5):
    raw = problem.case3d(40, 36, 32, 0.25, 0.01, 60, na=na, seed=seed, rmed=(0.05, 2.0, 0.5), src_depth=3.0,
                         nst_side=3, zbeg=zbeg, rmed_tile=(32, 32, 32), benchmark=benchmark, device_medium=True)
    m = raw.medium
    host = md.medium_setup3(raw.grid, m.rho, m.lam, m.mu, m.qp, m.qs, benchmark=benchmark)
    return raw, host


@pytest.mark.parametrize("na,zbeg,benchmark", [(8, -0.75, False), (2, -0.5, False), (8, -0.75, True),
                                               (6, 0.0, False)])
def test_device_setup_matches_host(na, zbeg, benchmark):
    raw, host = _pair(na, zbeg, benchmark)
    sim = problem.open_case3d(raw)
    try:
        got = {n: sim.medium_array(n) for n in EXACT + RELAXED}
    finally:
        sim.close()
    for n in EXACT:
        assert np.array_equal(got[n], getattr(host, n), equal_nan=True), n
    for n in RELAXED:
        want = getattr(host, n)
        err = np.abs(got[n] - want).max() / max(float(np.abs(want).max()), 1e-30)
        print(f"{n}: max |device - host| / max |host| = {err:.2e}")
        assert np.allclose(got[n], want, rtol=5e-6, atol=1e-6 * float(np.abs(want).max())), n
    m = raw.medium
    assert np.array_equal(m.kfs, host.kfs) and np.array_equal(m.kob, host.kob)
    assert m.vmin == pytest.approx(host.vmin, rel=1e-6)
    assert m.vmax == pytest.approx(host.vmax, rel=1e-6)
    if zbeg < 0:
        assert m.kfs[3:-3, 3:-3].min() > 0   # the air layer was detected


def test_device_setup_run_matches_host_setup():
    """Receiver traces of 60 steps: the device-set-up context against one
    built from the host assembly (same receivers)."""
    raw, host = _pair(8, -0.75, False, seed=9)
    dev = problem.open_case3d(raw)
    hc = problem.Case3D(raw.grid, host, raw.pml, raw.glut, raw.stations, raw.station_list, raw.nt, raw.ntdec_w,
                        raw.mech)
    ref = Swpc3D(hc.grid, hc.medium, hc.pml, hc.glut, hc.stations, hc.ntdec_w, hc.nt)
    try:
        for s in (dev, ref):
            s.run(1, raw.nt)
        a, b = dev.receivers()[0], ref.receivers()[0]
    finally:
        dev.close()
        ref.close()
    e = np.linalg.norm(a - b) / np.linalg.norm(b)
    print(f"trace rel-L2 device setup vs host setup: {e:.2e}")
    assert np.linalg.norm(b) > 0
    assert e < 1e-5


def test_device_setup_rejects_16bit_medium():
    raw, _ = _pair(8, -0.75, False)
    with pytest.raises(SwpcError, match="FP32 medium"):
        problem.open_case3d(raw, precision="fp32_s16")


def test_device_setup_without_absorber():
    """na = 0 (the learned-boundary grid): homogenisation only fills the
    memory box's halo from the outermost cells."""
    g = md.Grid3D(40, 36, 32, 0.25, 0.25, 0.25, 0.01, na=0, zbeg=-0.5)
    rho, lam, mu, qp, qs = md.layered3(g, problem.LHM_EXAMPLE)
    host = md.medium_setup3(g, rho, lam, mu, qp, qs)
    raw = md.RawMedium3D(rho, lam, mu, qp, qs)
    prof = tuple(np.zeros((n, 4), np.float32) for n in (40, 40, 36, 36, 32, 32))
    with Swpc3D(g, raw, prof, None, None, 1, 10) as sim:
        got = {n: sim.medium_array(n) for n in EXACT + RELAXED}
    for n in EXACT:
        assert np.array_equal(got[n], getattr(host, n)), n
    for n in RELAXED:
        want = getattr(host, n)
        assert np.allclose(got[n], want, rtol=5e-6, atol=1e-6 * float(np.abs(want).max())), n
    assert np.array_equal(raw.kfs, host.kfs) and np.array_equal(raw.kob, host.kob)


def test_device_setup_varying_q():
    """Q varying cell by cell: the relaxation factor is formed for ~10^5
    distinct tau values, FP64-then-rounded on the device against the host's
    complex FP32 steps."""
    raw, _ = _pair(8, -0.75, False, seed=11)
    m = raw.medium
    rng = np.random.default_rng(3)
    m.qp = (m.qp * rng.uniform(0.3, 3.0, m.qp.shape)).astype(np.float32)
    m.qs = (m.qs * rng.uniform(0.3, 3.0, m.qs.shape)).astype(np.float32)
    host = md.medium_setup3(raw.grid, m.rho, m.lam, m.mu, m.qp, m.qs)
    sim = problem.open_case3d(raw)
    try:
        got = {n: sim.medium_array(n) for n in EXACT + RELAXED}
    finally:
        sim.close()
    for n in EXACT:
        assert np.array_equal(got[n], getattr(host, n)), n
    nbits = 0
    for n in RELAXED:
        want = getattr(host, n)
        nbits += int((got[n] != want).sum())
        rel = np.abs(got[n] - want) / np.maximum(np.abs(want), 1e-30)
        print(f"{n}: {int((got[n] != want).sum())} cells differ, max rel {float(rel.max()):.2e}")
        assert np.allclose(got[n], want, rtol=5e-6, atol=1e-6 * float(np.abs(want).max())), n
    assert np.array_equal(m.kfs, host.kfs) and np.array_equal(m.kob, host.kob)
