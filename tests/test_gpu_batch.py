"""Batches of independent runs on the GPU: members advanced by one captured
graph per k steps give exactly the traces of their own runs, through a
graph replay, a remainder of single steps and a member reconfigured
between calls (which forces a recapture)."""
import numpy as np
import pytest

from openswpc_amd import batch, problem

pytestmark = pytest.mark.gpu


def _cases(nt):
    out = [("3d", problem.case3d(48, 40, 36, 0.25, 0.01, nt, na=8, seed=11, nst_side=2, src_depth=3.0, zbeg=-0.75,
                                 src_xy=(0.7, -0.4))),
           ("2d", problem.case2d(96, 64, 0.1, 0.005, nt, na=10, seed=12, nst=4, src_depth=2.0, src_x=0.35)),
           ("3d", problem.case3d(40, 44, 32, 0.25, 0.01, nt, na=8, seed=13, nst_side=2, src_depth=3.0, zbeg=-0.75,
                                 rmed=(0.05, 2.0, 0.5), rmed_tile=(32, 32, 32))),
           ("2d", problem.case2d(120, 80, 0.1, 0.005, nt, na=10, seed=14, nst=4, src_depth=3.0,
                                 rmed=(0.05, 2.0, 0.5)))]
    return out


def test_batch_equals_individual_runs():
    nt = 75
    cases = _cases(nt)
    ref = []
    for kind, c in cases:
        s = batch.open_case(kind, c)
        s.run(1, nt)
        ref.append(s.receivers()[0].copy())
        s.close()
    sims = [batch.open_case(kind, c) for kind, c in cases]
    with batch.Batch() as b:
        for s in sims:
            b.add(s)
        b.run(1, 40, 16)            # two graph replays + 8 single steps
        sims[0].set_tuning(4, 3, 0, 4, 5, 0, 32, 0, 1, 2, 0, 1, 2, 0, 8, 0)  # a setter: the graph is recaptured
        b.run(41, nt - 40, 16)
        b.sync()
    for (kind, c), s, r in zip(cases, sims, ref):
        v = s.receivers()[0]
        assert np.isfinite(v).all() and np.abs(v).max() > 0
        assert np.array_equal(v, r), kind
        s.close()


def test_batch_rejects_steps_beyond_the_source_table():
    cases = _cases(20)[:2]
    sims = [batch.open_case(kind, c) for kind, c in cases]
    with batch.Batch() as b:
        for s in sims:
            b.add(s)
        with pytest.raises(Exception, match="outside"):
            b.run(1, 21, 16)
        with pytest.raises(Exception, match="already"):
            b.add(sims[0])
    for s in sims:
        s.close()


def test_psv_group_equals_branches_and_own_runs():
    """The P-SV group (one launch per phase for every qualifying FP32 P-SV
    member) against the same batch with every member a graph branch and
    against each member's own run, bit for bit: five P-SV members, of which
    one has a snapshot stream and one runs in FP64 (both stay branches),
    one source on the first column (its stress glut stays a launch of its
    own, so the member stays a branch too) and one body force, beside a
    3-D member."""
    from openswpc_amd import snapshot
    from openswpc_amd.solver import SwpcPSV
    nt = 70
    cases = _cases(nt)
    g0 = problem.case2d(110, 70, 0.1, 0.005, nt, na=10, src_depth=2.0).grid
    extra = [("2d", problem.case2d(110, 70, 0.1, 0.005, nt, na=10, seed=15, nst=3, src_depth=2.0,
                                   src_x=g0.xbeg + 0.5 * g0.dx)),
             ("2d", problem.case2d(100, 72, 0.1, 0.005, nt, na=10, seed=16, nst=3, src_depth=2.5, bodyforce=True)),
             ("2d", problem.case2d(90, 66, 0.1, 0.005, nt, na=10, seed=17, nst=3, src_depth=2.0))]
    cases = cases + extra
    prec = ["fp32"] * len(cases)
    prec[-1] = "fp64"
    snap_member = 1

    def open_all():
        sims = []
        for q, (kind, c) in enumerate(cases):
            if kind == "3d":
                sims.append(batch.open_case(kind, c))
            else:
                sims.append(SwpcPSV(c.grid, c.medium, c.pml, c.glut, c.stations, c.ntdec_w, nt, precision=prec[q],
                                    bodyforce=c.bforce))
        c = cases[snap_member][1]
        _, _, ik = snapshot.cells2(snapshot.snap_grid2(c.grid, 4, 4))
        sid = sims[snap_member].add_snapshot(1, ik, 7, 16, snapshot.UC2, 1.0, np.float32(c.grid.dt))
        return sims, sid, len(ik)

    out = {}
    for mode in ("own", "group", "branches"):
        sims, sid, npts = open_all()
        if mode == "own":
            for s in sims:
                s.run(1, nt)
        else:
            with batch.Batch() as b:
                b.set_grouping(psv=mode == "group", d3=False)
                for s in sims:
                    b.add(s)
                b.run(1, nt, 16)
                b.sync()
        out[mode] = [s.receivers()[0].copy() for s in sims] + [sims[snap_member].get_snapshot(sid, 9, 2, npts)]
        for s in sims:
            s.close()
    for q, ref in enumerate(out["own"]):
        assert np.isfinite(ref).all() and np.abs(ref).max() > 0, q
        assert np.array_equal(out["group"][q], ref), ("group", q)
        assert np.array_equal(out["branches"][q], ref), ("branches", q)


def test_3d_group_equals_branches_and_own_runs():
    """The 3-D group (one launch per phase for every qualifying FP32 3-D
    member) against the same batch with every member a graph branch and
    against each member's own run, bit for bit, receivers and every field:
    four FP32 3-D members of uneven size (a ragged one with a random medium
    set up on the GPU, one 300 cells wide whose right x face lies past its
    first interior tile, one retuned to other launch shapes, which the
    group does not use), one FP64 member and one with a body force (both
    stay branches), beside a P-SV member."""
    from openswpc_amd.solver import FIELDS3, Swpc3D
    nt = 70
    base = _cases(nt)
    cases = [base[0], base[1], base[2],
             ("3d", problem.case3d(71, 45, 39, 0.25, 0.01, nt, na=6, seed=21, nst_side=2, src_depth=2.5, zbeg=-0.75,
                                   rmed=(0.05, 2.0, 0.5), rmed_tile=(32, 32, 32), device_medium=True)),
             ("3d", problem.case3d(300, 40, 36, 0.25, 0.01, nt, na=8, seed=22, nst_side=2, src_depth=3.0, zbeg=-0.75,
                                   src_xy=(33.0, 0.3), st_spacing=4.0)),
             ("3d", problem.case3d(44, 40, 34, 0.25, 0.01, nt, na=8, seed=23, nst_side=2, src_depth=3.0, zbeg=-0.75)),
             ("3d", problem.case3d(46, 42, 34, 0.25, 0.01, nt, na=8, seed=24, nst_side=2, src_depth=3.0, zbeg=-0.75,
                                   bodyforce=True))]
    fp64_member, retuned = 5, 2

    def open_all():
        sims = []
        for q, (kind, c) in enumerate(cases):
            if q == fp64_member:
                sims.append(Swpc3D(c.grid, c.medium, c.pml, c.glut, c.stations, c.ntdec_w, nt, precision="fp64"))
            else:
                sims.append(batch.open_case(kind, c))
        # entries 0..15 of swpc3d_set_tuning: 256 x 2 staged stress, velocity tile rows 8, chunk 16
        sims[retuned].set_tuning(4, 3, 0, 4, 5, 0, 32, 0, 1, 2, 0, 1, 8, 0, 16, 0)
        return sims

    fields = ("Vx", "Vy", "Vz", "Sxx", "Syz", "Sxy", "Rxx1", "Rxz3", "axVx", "ayVy", "azSzz", "axSxy")
    assert set(fields) <= set(FIELDS3)
    out = {}
    for mode in ("own", "group", "branches"):
        sims = open_all()
        if mode == "own":
            for s in sims:
                s.run(1, nt)
        else:
            with batch.Batch() as b:
                b.set_grouping(psv=mode == "group", d3=mode == "group")
                for s in sims:
                    b.add(s)
                b.run(1, nt, 16)
                b.sync()
        res = []
        for (kind, _), s in zip(cases, sims):
            res.append(s.receivers()[0].copy())
            if kind == "3d":
                res += [s.field(f) for f in fields]
        out[mode] = res
        for s in sims:
            s.close()
    assert len(out["own"]) == len(out["group"]) == len(out["branches"])
    for q, ref in enumerate(out["own"]):
        assert np.isfinite(ref).all(), q
        assert np.array_equal(out["group"][q], ref), ("group", q)
        assert np.array_equal(out["branches"][q], ref), ("branches", q)
    # every member moved (the wide member's source sits 2.5 km from its
    # right absorber, far from its receivers: its fields carry the signal)
    q = 0
    for kind, _ in cases:
        assert np.abs(out["own"][q + (1 if kind == "3d" else 0)]).max() > 0, q
        q += 1 + (len(fields) if kind == "3d" else 0)
