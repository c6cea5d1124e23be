"""Captured step graphs (the step number in device memory, graph_k steps
per replay) against step-by-step launches: bit-identical receivers,
snapshots and fields, with eager work (field reads, vmax, single steps)
between replays and runs that end in a partial graph.
"""
import numpy as np
import pytest

from openswpc_amd import problem, snapshot
from openswpc_amd.solver import Swpc3D, SwpcPSV

pytestmark = pytest.mark.gpu
F32 = np.float32


def run3(case, graph_k, prec):
    g = case.grid
    sg = snapshot.snap_grid3(g, 2, 2, 2)
    with Swpc3D(g, case.medium, case.pml, case.glut, case.stations, 2, case.nt, precision=prec) as sim:
        # the measured defaults of entries 0..15; entry 16 is the graph length
        tun = [4, 3, 0, 4, 5, 0, 32, 0, 1, 2, 0, 1 if prec != "fp32_s16" else 0, 4, 0, 8, 0, graph_k]
        sim.set_tuning(*tun)
        _, _, ijk = snapshot.cells3(sg, "xy", case.medium.kfs, case.medium.kob)
        sids = [sim.add_snapshot(k, ijk, 5, 64, snapshot.UC3, F32(case.glut.M0), F32(g.dt)) for k in range(3)]
        out = []
        it = 1
        for n in (37, 1, 16, 33, 3):       # partial graphs, single steps, exact multiples
            sim.run(it, n)
            it += n
            out.append(sim.field("Vx"))
            out.append(sim.kernel__vmax())
        vel, disp = sim.receivers()
        frames = [[sim.get_snapshot(s, f, 4 if k == 0 else 3, len(ijk)) for f in range((it - 2) // 5 + 1)]
                  for k, s in enumerate(sids)]
        return vel, disp, out, frames, sim.field("Sxz"), sim.field("azSzz")


@pytest.mark.parametrize("prec", ["fp32", "fp64", "fp32_r16"])
def test_3d_graph_equals_step_by_step(prec):
    print(f"3-D graphs vs step by step ({prec})", flush=True)
    case = problem.case3d(64, 56, 48, 0.5, 0.04, 90, na=8, src_depth=5.0, nst_side=3, zbeg=-2.0)
    a = run3(case, 16, prec)
    b = run3(case, 0, prec)
    assert np.abs(a[0]).max() > 0
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    for x, y in zip(a[2], b[2]):
        assert np.array_equal(x, y)
    for fa, fb in zip(a[3], b[3]):
        for x, y in zip(fa, fb):
            assert np.array_equal(x, y)
    assert np.array_equal(a[4], b[4]) and np.array_equal(a[5], b[5])


def run2(case, graph_k):
    g = case.grid
    sg = snapshot.snap_grid2(g, 2, 2)
    _, _, ik = snapshot.cells2(sg)
    with SwpcPSV(g, case.medium, case.pml, case.glut, case.stations, 3, case.nt) as sim:
        sim.set_graph(graph_k)
        sid = sim.add_snapshot(2, ik, 7, 64, snapshot.UC2, F32(case.glut.M0), F32(g.dt))
        it = 1
        mids = []
        for n in (50, 17, 1, 64, 8):
            sim.run(it, n)
            it += n
            mids.append(sim.field("Vz"))
        vel, disp = sim.receivers()
        fr = [sim.get_snapshot(sid, f, 2, len(ik)) for f in range((it - 2) // 7 + 1)]
        return vel, disp, mids, fr


def test_psv_graph_equals_step_by_step():
    print("P-SV graphs vs step by step", flush=True)
    case = problem.case2d(200, 100, 0.25, 0.02, 140, na=10, src_depth=5.0, zbeg=-2.0)
    a = run2(case, 16)
    b = run2(case, 0)
    c = run2(case, 5)
    assert np.abs(a[0]).max() > 0
    for other in (b, c):
        assert np.array_equal(a[0], other[0]) and np.array_equal(a[1], other[1])
        for x, y in zip(a[2], other[2]):
            assert np.array_equal(x, y)
        for x, y in zip(a[3], other[3]):
            assert np.array_equal(x, y)
