"""The Fortran side of the boundary: include/swpc_hip.f90 (bind(C)
interfaces) and the drop-in host openswpc_amd/fhost (reference-named
kernel__update_stress, source__stressglut, wav__store, ... in the order of
src/swpc_3d/main.f90:125-144), built with amdflang.

CPU: the bindings pass descriptors, values and error text correctly.
GPU: the Fortran host run equals the Python host run through the same
library, bit for bit, and the FP32 oracle within the kernels' bar.
"""
import subprocess

import numpy as np
import pytest

from openswpc_amd import build, fcase, problem


@pytest.fixture(scope="module")
def fdrive():
    exe = build.FDRIVE
    if not exe.exists():
        if build.build_fortran_host() is None:
            pytest.skip("no Fortran compiler to build the Fortran host")
    return exe


def test_fortran_bindings_without_device(fdrive):
    r = subprocess.run([str(fdrive), "check"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bindings ok: grid 4x64x64 too small" in r.stdout


@pytest.mark.gpu
def test_fortran_host_matches_python_host(fdrive, tmp_path):
    from openswpc_amd.solver import Swpc3D
    from oracle.oracle import Oracle3D
    nt = 50
    case = problem.case3d(60, 52, 44, 0.5, 0.04, nt, na=8, src_depth=5.0, nst_side=3, zbeg=-2.0)
    print("Fortran host: 60x52x44, nt=50", flush=True)
    ntw, nst = fcase.write_case3(tmp_path / "case.bin", case)
    r = subprocess.run([str(fdrive), str(tmp_path / "case.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    vel, disp, vmax = fcase.read_result3(tmp_path / "out.bin", ntw, nst)
    with Swpc3D(case.grid, case.medium, case.pml, case.glut, case.stations, case.ntdec_w, nt) as sim:
        sim.run(1, nt)
        v2, d2 = sim.receivers()
        vm2 = sim.kernel__vmax()
    assert np.abs(v2).max() > 0
    assert np.array_equal(vel, v2) and np.array_equal(disp, d2) and np.array_equal(vmax, vm2)
    ref = Oracle3D(case, 32)
    ref.run(1, nt)
    assert np.linalg.norm(vel - ref.vel) / np.linalg.norm(ref.vel) <= 1e-5
