"""Autotuner for the interior launch configuration of the 3-D kernels.

Times candidate tile shapes / register budgets / z-chunk lengths on the
actual case with HIP events recorded on the context's stream (a few steps
per candidate) and installs the fastest with ``Swpc3D.set_tuning``.  The
results of the kernels do not depend on the configuration.
"""
from __future__ import annotations

import itertools

import numpy as np

# (by, wpe, split) stress candidates; (by, wpe) velocity candidates
STRESS = [(4, 0, 0), (2, 0, 0), (8, 0, 0), (4, 3, 0), (4, 4, 0), (2, 4, 0), (8, 4, 0),
          (4, 0, 1), (2, 0, 1), (8, 0, 1), (4, 4, 1), (2, 4, 1), (8, 4, 1)]
VEL = [(4, 0), (2, 0), (8, 0), (4, 5), (8, 5)]


def _time(sim, it0, steps, cls):
    sim.sync()
    sim.profile(True)
    sim.run(it0, steps)
    ms, cnt = sim.profile_read()
    sim.profile(False)
    return ms[cls] / max(cnt[cls], 1)


def tune(sim, it0=1, steps=6, kcs=(0,), xcds=(1, 0), verbose=False):
    """Returns (best tuning vector, table of (config, ms)); leaves the best
    installed.  Advances the simulation by the steps it times."""
    it = it0
    table = []
    best_s, best_v = None, None
    for xcd in xcds:
        for (by, wpe, split), kc in itertools.product(STRESS, kcs):
            sim.set_tuning(by, wpe, split, 4, 0, xcd, kc, 0)
            t = _time(sim, it, steps, 0)
            it += steps
            table.append((("stress", by, wpe, split, xcd, kc), t))
            if verbose:
                print(f"stress by={by} wpe={wpe} split={split} xcd={xcd} kc={kc}: {t * 1e3:.1f} us", flush=True)
            if best_s is None or t < best_s[1]:
                best_s = ((by, wpe, split, xcd, kc), t)
    by, wpe, split, xcd, kc = best_s[0]
    for (vby, vwpe), vkc in itertools.product(VEL, kcs):
        sim.set_tuning(by, wpe, split, vby, vwpe, xcd, kc, vkc)
        t = _time(sim, it, steps, 2)
        it += steps
        table.append((("vel", vby, vwpe, xcd, vkc), t))
        if verbose:
            print(f"vel by={vby} wpe={vwpe} xcd={xcd} kc={vkc}: {t * 1e3:.1f} us", flush=True)
        if best_v is None or t < best_v[1]:
            best_v = ((vby, vwpe, vkc), t)
    vby, vwpe, vkc = best_v[0]
    best = (by, wpe, split, vby, vwpe, xcd, kc, vkc)
    sim.set_tuning(*best)
    return best, table, it
