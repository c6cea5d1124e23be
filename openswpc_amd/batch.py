"""Batches of independent runs on one GPU (no counterpart in the reference,
which runs one simulation per MPI job).

``Batch`` advances many 3-D and P-SV contexts of uneven size together:
graph_k steps of every member are captured once into one HIP graph, one
branch per member, and replayed with one launch per graph_k steps
(swpc_batch_run).  Each member keeps its own stream, step counter and
state, so its results equal those of its own run.

``mixed_cases`` builds the mixed workload: P-SV grids of 256-1024 cells
per side and 3-D grids of 96-256 cells per side, sizes, source positions
and mechanisms drawn from the seed, layered media with a von Karman
perturbation.  ``tune_members`` picks the 3-D kernels' launch shapes per
size class with the autotuner on a scratch context.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import problem
from ._lib import check, lib


class Batch:
    def __init__(self):
        h = C.c_void_p()
        check(lib().swpc_batch_create(C.byref(h)), "swpc_batch_create")
        self.h = h
        self.members = []

    def add(self, sim):
        """Add a Swpc3D or SwpcPSV; it must stay open while the batch lives."""
        from .solver import Swpc3D
        f = lib().swpc_batch_add3d if isinstance(sim, Swpc3D) else lib().swpc_batch_add2d
        check(f(self.h, sim.h), "swpc_batch_add")
        self.members.append(sim)

    def run(self, it0, nsteps, graph_k=16):
        check(lib().swpc_batch_run(self.h, it0, nsteps, graph_k), "swpc_batch_run")

    def set_grouping(self, psv=True, d3=True):
        """Qualifying FP32 P-SV members (psv) and FP32 3-D members (d3) in
        groups, one launch per phase for all members of a group, instead of
        one graph branch each (both by default); results are the same."""
        check(lib().swpc_batch_set_grouping(self.h, int(bool(psv)) | (2 if d3 else 0)), "swpc_batch_set_grouping")

    def sync(self):
        check(lib().swpc_batch_sync(self.h), "swpc_batch_sync")

    def close(self):
        if getattr(self, "h", None):
            lib().swpc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def mixed_cases(seed=1, n2d=24, n3d=24, nt=2000, side2=(256, 1024), side3=(96, 256), na=20, eps=0.05):
    """The mixed batch: n2d P-SV runs (dx = 0.1 km, dt = 5 ms, the two-layer
    model) and n3d 3-D runs (dx = 0.25 km, dt = 10 ms, the example's
    layered model), each with a von Karman perturbation (eps, a = 2 km,
    kappa = 0.5), a double couple at a random position inside the
    interior and a random mechanism, and a few surface receivers.  Yields
    ('2d'|'3d', case) in the seed's order, built lazily; the 3-D cases
    leave medium__setup to the GPU (problem.case3d(device_medium=True))."""
    for kind, kw in mixed_specs(seed, n2d, n3d, nt, side2, side3, na, eps):
        yield kind, build_case(kind, kw)


def build_case(kind, kw):
    """One member of mixed_specs; kw may override device_medium (False:
    the 3-D medium assembled on the host, as the oracle needs it)."""
    return problem.case2d(**kw) if kind == "2d" else problem.case3d(**kw)


def mixed_specs(seed=1, n2d=24, n3d=24, nt=2000, side2=(256, 1024), side3=(96, 256), na=20, eps=0.05):
    """The (kind, case-builder arguments) list behind mixed_cases."""
    rng = np.random.default_rng(seed)
    specs = []
    for q in range(n2d + n3d):
        s = int(seed) * 7919 + q
        if q < n2d:
            nx, nz = (int(v) for v in rng.integers(side2[0], side2[1] + 1, 2))
            dx = 0.1
            hx = (nx / 2 - na - 5) * dx
            depth = float(rng.uniform(2.0, min(20.0, (nz - na - 8) * dx - 0.75)))
            specs.append(("2d", dict(nx=nx, nz=nz, dx=dx, dt=0.005, nt=nt, na=na, seed=s, rmed=(eps, 2.0, 0.5),
                                     src_depth=depth, nst=4, st_spacing=2.0, zbeg=-0.3,
                                     src_x=float(rng.uniform(-hx, hx)))))
        else:
            nx, ny, nz = (int(v) for v in rng.integers(side3[0], side3[1] + 1, 3))
            dx = 0.25
            hx, hy = (nx / 2 - na - 5) * dx, (ny / 2 - na - 5) * dx
            depth = float(rng.uniform(2.0, min(20.0, (nz - na - 8) * dx - 0.75)))
            specs.append(("3d", dict(nx=nx, ny=ny, nz=nz, dx=dx, dt=0.01, nt=nt, na=na, seed=s,
                                     rmed=(eps, 2.0, 0.5), src_depth=depth, nst_side=2, zbeg=-0.75,
                                     rmed_tile=(64, 64, 64), device_medium=True,
                                     src_xy=(float(rng.uniform(-hx, hx)), float(rng.uniform(-hy, hy))))))
    return specs


def open_case(kind, c, precision="fp32"):
    """A context for one case; 3-D cases set their medium up on the GPU."""
    from .solver import SwpcPSV
    if kind == "3d":
        return problem.open_case3d(c, precision)
    return SwpcPSV(c.grid, c.medium, c.pml, c.glut, c.stations, c.ntdec_w, c.nt, precision=precision)


def size_class(g):
    """3-D grids tuned together: sides rounded to 32 cells."""
    return tuple(int(round(n / 32)) for n in (g.nx, g.ny, g.nz))


def tune_members(cases, steps=3, log=None):
    """Autotuned launch shapes per 3-D size class, each on a scratch context
    of the first case of its class: {class: tuning vector}."""
    from . import autotune
    out = {}
    for kind, c in cases:
        if kind != "3d":
            continue
        k = size_class(c.grid)
        if k in out:
            continue
        sim = open_case(kind, c)
        best, table, _ = autotune.tune(sim, 1, steps)
        sim.close()
        out[k] = best
        if log:
            log(f"batch: tuned {c.grid.nx}x{c.grid.ny}x{c.grid.nz}: {best}")
    return out


def cells(cases):
    n = 0
    for kind, c in cases:
        g = c.grid
        n += g.nx * g.nz * (g.ny if kind == "3d" else 1)
    return n


def alg_bytes_per_step(cases):
    """Algorithmic HBM bytes of one step of every case, FP32: P-SV 152 B per
    interior and 136 B per absorber cell, 3-D 292 B and 284 B (the stress
    and velocity updates' compulsory traffic, m_kernel.f90 and
    m_absorb_p.f90 under the two-phase order)."""
    n = 0
    for kind, c in cases:
        g = c.grid
        if kind == "3d":
            ni = g.n_interior_cells()
            n += 292 * ni + 284 * (g.nx * g.ny * g.nz - ni)
        else:
            ni = max(g.nx - 2 * g.na, 0) * max(g.nz - g.na, 0)
            n += 152 * ni + 136 * (g.nx * g.nz - ni)
    return n


def measure(cases, steps, graph_k=16, tuning=None, log=None):
    """Time `steps` steps of every case five ways on fresh contexts:
    'batch' (one graph per graph_k steps of all members, the P-SV members
    in one group and the 3-D members in another, one launch per phase for
    all members of a group; the 3-D group uses launch shapes of its own),
    'batch_psv_group' (the P-SV group only, the 3-D members autotuned
    branches), 'branches' (the
    same graph with every member a branch of its own), 'streams' (each
    member's own run, all enqueued before one wait: the members overlap on
    their own streams) and 'sequential' (one member after the other).
    Returns {mode: seconds} plus the total cells."""
    res = {}
    for mode in ("batch", "batch_psv_group", "branches", "streams", "sequential"):
        sims = []
        for kind, c in cases:
            s = open_case(kind, c)
            # the 3-D group uses launch shapes of its own
            if tuning and kind == "3d" and size_class(c.grid) in tuning and mode != "batch":
                s.set_tuning(*tuning[size_class(c.grid)])
            sims.append(s)
        warm = min(graph_k or 16, steps)
        if mode in ("batch", "batch_psv_group", "branches"):
            b = Batch()
            b.set_grouping(psv=mode != "branches", d3=mode == "batch")
            for s in sims:
                b.add(s)
            b.run(1, warm, graph_k)
            b.sync()
            t0 = time.perf_counter()
            b.run(warm + 1, steps - warm, graph_k)
            b.sync()
            dt = time.perf_counter() - t0
            b.close()
        else:
            for s in sims:
                s.run(1, warm)
            for s in sims:
                s.sync()
            t0 = time.perf_counter()
            for s in sims:
                s.run(warm + 1, steps - warm)
                if mode == "sequential":
                    s.sync()
            for s in sims:
                s.sync()
            dt = time.perf_counter() - t0
        res[mode] = dt
        if log:
            log(f"batch: {mode}: {steps - warm} steps of {len(sims)} runs in {dt:.3f} s")
        for s in sims:
            s.close()
    res["cells"] = cells(cases)
    res["timed_steps"] = steps - min(graph_k or 16, steps)
    return res


def _main(argv=None):
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(prog="python -m openswpc_amd.batch",
                                 description="time the mixed batch of independent runs on one GPU")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--n2d", type=int, default=24)
    ap.add_argument("--n3d", type=int, default=24)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--graph-k", type=int, default=16)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--side2", type=int, nargs=2, default=(256, 1024), help="P-SV side range (cells)")
    ap.add_argument("--side3", type=int, nargs=2, default=(96, 256), help="3-D side range (cells)")
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.time()
    cases = list(mixed_cases(a.seed, a.n2d, a.n3d, a.steps, side2=tuple(a.side2), side3=tuple(a.side3)))
    log(f"batch: {len(cases)} cases built in {time.time() - t0:.1f} s")
    tuning = None if a.no_tune else tune_members(cases, log=log)
    res = measure(cases, a.steps, a.graph_k, tuning, log)
    n = res["cells"] * res["timed_steps"]
    ab = alg_bytes_per_step(cases) * res["timed_steps"]
    out = {"workload": f"{a.n2d} P-SV ({a.side2[0]}-{a.side2[1]} per side) + {a.n3d} 3-D ({a.side3[0]}-{a.side3[1]} "
                       f"per side) runs, seed {a.seed}, {a.steps} steps each",
           "cells": res["cells"], "timed_steps": res["timed_steps"], "graph_k": a.graph_k,
           "tuned_classes": len(tuning or {}),
           **{f"{m}_cell_updates_per_s": n / res[m]
              for m in ("batch", "batch_psv_group", "branches", "streams", "sequential")},
           "speedup_vs_sequential": res["sequential"] / res["batch"],
           "speedup_vs_branches": res["branches"] / res["batch"],
           "batch_psv_group_vs_branches": res["branches"] / res["batch_psv_group"],
           "alg_bytes_per_step": alg_bytes_per_step(cases),
           "batch_alg_GBs": ab / res["batch"] / 1e9, "batch_hbm_frac": ab / res["batch"] / 8e12}
    print(json.dumps(out))


if __name__ == "__main__":
    _main()
