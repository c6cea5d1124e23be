"""report__progress logic (src/swpc_3d/m_report.f90:120-185) of
run_monitored, with a stand-in solver on the CPU: where the check falls in
the time loop, the divergence threshold 1e5 on max|V|*UC, non-finite
amplitudes, and the report line."""
import math

import numpy as np
import pytest

from openswpc_amd.solver import DIVERGENCE_TOL, NumericalDivergence, _run_monitored


class FakeSim:
    UC = 1e-15
    M0 = 2e15

    def __init__(self, blow_at=None, value=1.0, start=1):
        self.steps = []
        self.checks = []
        self.blow_at = blow_at
        self.value = value
        self.cur = start          # the next step to run

    def run(self, it0, n):
        assert it0 == self.cur
        self.steps.extend(range(it0, it0 + n))
        self.cur = it0 + n

    def kernel__vmax(self):
        it = self.cur
        self.checks.append(it)
        if self.blow_at is not None and it >= self.blow_at:
            return np.array([self.value, 0.0, 0.0], np.float32)
        return np.array([1.0, 2.0, 3.0], np.float32)


def test_checks_fall_on_multiples_before_the_step():
    sim = FakeSim()
    lines = []
    _run_monitored(sim, 1, 350, 100, lines.append)
    assert sim.steps == list(range(1, 351))
    assert sim.checks == [100, 200, 300]
    assert len(lines) == 3 and lines[0].startswith("  it=0000100,")
    # physical units: vmax * UC * M0
    assert "2.00E+00" in lines[0] and "6.00E+00" in lines[0]
    sim = FakeSim(start=200)
    _run_monitored(sim, 200, 101, 100, None)
    assert sim.checks == [200, 300] and sim.steps == list(range(200, 301))


def test_divergence_threshold_and_nan():
    sim = FakeSim(blow_at=300, value=np.float32(DIVERGENCE_TOL / FakeSim.UC * 1.01))
    with pytest.raises(NumericalDivergence, match="it=300"):
        _run_monitored(sim, 1, 1000, 100, None)
    assert sim.steps[-1] == 299
    sim = FakeSim(blow_at=200, value=np.float32(DIVERGENCE_TOL / FakeSim.UC * 0.99))
    _run_monitored(sim, 1, 400, 100, None)       # below the threshold: no abort
    sim = FakeSim(blow_at=200, value=np.float32(math.nan))
    with pytest.raises(NumericalDivergence):
        _run_monitored(sim, 1, 400, 100, None)
