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
