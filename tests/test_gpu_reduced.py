"""Reduced-storage forms against the FP64 run of the same scheme: FP32
arithmetic; memory variables in FP16 with a per-array power-of-two scale;
the medium as an FP32 per-plane base plus an FP16 residual and the ADE
arrays in BF16 (fp32_s16), or both in FP32 (fp32_r16).  Bar: seismogram
relative L2 misfit <= 1e-3.
"""
import numpy as np
import pytest

from openswpc_amd import problem
from openswpc_amd.solver import Swpc3D

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def run(case, precision, nt):
    with Swpc3D(case.grid, case.medium, case.pml, case.glut, case.stations, case.ntdec_w, nt,
                precision=precision) as sim:
        sim.run(1, nt)
        vel, disp = sim.receivers()
        return vel, disp, sim.field("Vz")


@pytest.mark.parametrize("form", ["fp32_s16", "fp32_r16"])
@pytest.mark.parametrize("rmed", [None, (0.05, 2.0, 0.5)])
def test_reduced_storage_misfit_vs_fp64(rmed, form):
    nt = 400
    case = problem.case3d(128, 120, 80, 0.25, 0.01, nt, na=12, src_depth=5.0, nst_side=4, zbeg=-2.0,
                          rmed=rmed, rmed_tile=(64, 64, 64))
    print(f"reduced ({form}) vs fp64: 128x120x80 nt={nt} rmed={rmed}", flush=True)
    v64, d64, vz64 = run(case, "fp64", nt)
    v16, d16, vz16 = run(case, form, nt)
    v32, d32, _ = run(case, "fp32", nt)
    e16 = rel(v16, v64)
    e32 = rel(v32, v64)
    print(f"misfit vs fp64: fp32 {e32:.2e}, reduced {e16:.2e}; disp {rel(d16, d64):.2e}; "
          f"wavefield {rel(vz16, vz64):.2e}", flush=True)
    assert np.abs(v64).max() > 0
    assert e32 < 1e-5
    assert e16 < 1e-3
    assert rel(d16, d64) < 1e-3


def test_reduced_storage_set_get_roundtrip():
    """Memory variables and ADE arrays survive a set/get through their
    16-bit storage within the format's rounding."""
    # the memory-variable scale follows the source's whole moment: give it
    # the full rise (1 s) so that values of the size set here are in range
    case = problem.case3d(48, 44, 40, 0.5, 0.04, 40, na=6, src_depth=4.0, nst_side=2, zbeg=-2.0)
    rng = np.random.default_rng(5)
    with Swpc3D(case.grid, case.medium, case.pml, case.glut, case.stations, 1, 40, precision="fp32_s16") as sim:
        a = rng.standard_normal(case.grid.shape) * 1e-3
        sim.set_field("Rxx2", a)
        b = sim.field("Rxx2")
        assert np.allclose(b, a, rtol=2e-3, atol=1e-3 * np.abs(a).max())
        sim.set_field("azSzz", a)
        b = sim.field("azSzz")
        assert np.allclose(b, a, rtol=8e-3, atol=1e-30)
