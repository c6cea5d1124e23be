"""Autotuner for the launch configuration of the 3-D kernels.

Times candidate kernel forms on the actual case with HIP events recorded
on the context's streams (a few steps per candidate) and installs the
fastest with ``Swpc3D.set_tuning``.

* FP32 and reduced storage use the 4-cell-per-thread kernels: candidates
  are the stress form (two launches / one pass / one pass z-marching with a
  chunk length) and the velocity z-marching chunk.
* FP64 uses the one-cell-per-thread kernels: candidates are tile rows,
  register budgets and the normal/shear split.

Results do not depend on the configuration beyond FP32 rounding order.
"""
from __future__ import annotations

import itertools

# one-cell-per-thread candidates: (by, wpe, split) for stress, (by, wpe) for velocity
STRESS = [(4, 0, 0), (2, 0, 0), (8, 0, 0), (4, 3, 0), (4, 4, 0), (2, 4, 0), (8, 4, 0),
          (4, 0, 1), (2, 0, 1), (8, 0, 1), (4, 4, 1), (2, 4, 1), (8, 4, 1)]
VEL = [(4, 0), (2, 0), (8, 0), (4, 5), (8, 5)]
# 4-wide candidates: (form, z-march chunk) for stress, z-march chunk for velocity
STRESS4 = [(0, 0), (1, 0), (2, 0), (1, 4), (1, 8), (1, 16)]
VEL4 = [0, 8, 16, 32]
# tuning vector defaults of the first eleven entries (see swpc3d_set_tuning;
# entry 5 = 2: the z-marching velocity update's y tiles in pairs per XCD)
BASE = [4, 3, 0, 4, 5, 2, 32, 0, 1, 2, 0]


def _time(sim, it0, steps, cls):
    sim.sync()
    sim.profile(True)
    sim.run(it0, steps)
    ms, cnt = sim.profile_read()
    sim.profile(False)
    return ms[cls] / max(cnt[cls], 1)


def tune(sim, it0=1, steps=6, verbose=False):
    """Returns (best tuning vector, table of (config, ms), next step);
    leaves the best installed.  Advances the simulation by the steps it
    times."""
    if sim.precision == "fp64":
        return _tune_1cell(sim, it0, steps, verbose)
    it = it0
    table = []
    best_s = best_v = None
    for form, sz in STRESS4:
        v = BASE + [form, 2, 0, 8, sz]
        sim.set_tuning(*v)
        t = _time(sim, it, steps, 0)
        it += steps
        table.append((("stress4", form, sz), t))
        if verbose:
            print(f"stress form={form} zchunk={sz}: {t * 1e3:.1f} us", flush=True)
        if best_s is None or t < best_s[1]:
            best_s = ((form, sz), t)
    form, sz = best_s[0]
    for vz in VEL4:
        sim.set_tuning(*(BASE + [form, 2, 0, vz, sz]))
        t = _time(sim, it, steps, 2)
        it += steps
        table.append((("vel4", vz), t))
        if verbose:
            print(f"vel zchunk={vz}: {t * 1e3:.1f} us", flush=True)
        if best_v is None or t < best_v[1]:
            best_v = (vz, t)
    best = BASE + [form, 2, 0, best_v[0], sz]
    sim.set_tuning(*best)
    return best, table, it


def _tune_1cell(sim, it0, steps, verbose):
    it = it0
    table = []
    best_s = best_v = None
    for xcd in (1, 0):
        for by, wpe, split in STRESS:
            sim.set_tuning(by, wpe, split, 4, 0, xcd, 0, 0)
            t = _time(sim, it, steps, 0)
            it += steps
            table.append((("stress", by, wpe, split, xcd), t))
            if verbose:
                print(f"stress by={by} wpe={wpe} split={split} xcd={xcd}: {t * 1e3:.1f} us", flush=True)
            if best_s is None or t < best_s[1]:
                best_s = ((by, wpe, split, xcd), t)
    by, wpe, split, xcd = best_s[0]
    for vby, vwpe in VEL:
        sim.set_tuning(by, wpe, split, vby, vwpe, xcd, 0, 0)
        t = _time(sim, it, steps, 2)
        it += steps
        table.append((("vel", vby, vwpe, xcd), t))
        if verbose:
            print(f"vel by={vby} wpe={vwpe} xcd={xcd}: {t * 1e3:.1f} us", flush=True)
        if best_v is None or t < best_v[1]:
            best_v = ((vby, vwpe), t)
    vby, vwpe = best_v[0]
    best = [by, wpe, split, vby, vwpe, xcd, 0, 0]
    sim.set_tuning(*best)
    return best, table, it
