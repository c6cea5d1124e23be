"""Benchmark of the swpc_3d time-stepping core on one MI355X.

Workload: the plain 3-D viscoelastic method at 512 x 512 x 256 cells,
dx = 0.25 km, dt = 10 ms, FP32; a 1-D layered medium (the reference's
example lhm.dat layers, Qp 600 / Qs 300, 3 Zener bodies) with a 5 %-rms
von Karman perturbation (a = 2 km, kappa = 0.5) from the seed, 20-cell
ADE-PML, one double-couple Kupper source at 10 km depth, 64 receivers on
the free surface.  A step is one full time step (stress + velocity
half-steps, PML, source, receivers) over the whole grid; inputs are
resident in HBM before the timed region.

Prints one JSON line with cell-updates/s, the HBM roofline of the dominant
kernel (HIP events on the library's own stream) and a CPU baseline: the
oracle (C restatement of the reference loops, OpenMP, all host cores) on a
bounded sample of the same grid.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s
# algorithmic bytes per cell-update (compulsory traffic, FP32, NM = 3)
B_INT_STRESS = 58 * 4   # V 3r, medium 7r, R 18r/18w, S 6r/6w
B_INT_STRESS_NEED = 55 * 4   # what k3x_stress streams: muyz, muxz, muxy come from mu and its neighbours
B_INT_VEL = 15 * 4      # S 6r, b 3r, V 3r/3w
B_PML_STRESS = 38 * 4
B_PML_VEL = 33 * 4
CHOSEN = "fp32_r16"    # the reduced-storage form reported beside the plain one
MISFIT_BAR = 1e-3      # receiver-trace misfit against the FP64 plain scheme that a faster form must meet


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-steps", type=int, default=60,
                    help="oracle steps of the CPU baseline (after one untimed step); the GPU fields are compared "
                         "with the oracle's after them, so the source (onset 0.1 s) must have started")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-chosen", action="store_true")
    ap.add_argument("--misfit-steps", type=int, default=2000,
                    help="steps of the misfit runs, independent of --steps: long enough for the P and S "
                         "arrivals from the 10 km deep source to cross the surface receivers and for the "
                         "absorbers' reflections to reach them (from about 8 s, 800 steps)")
    ap.add_argument("--no-c1", action="store_true", help="skip the P-SV 800x400 4000-step leg")
    ap.add_argument("--no-misfit", action="store_true",
                    help="skip the misfit runs (profiling: the timed forms only, no accuracy and no choice)")
    ap.add_argument("--lab-sponge", type=float, default=None,
                    help="c5: strip sponge strength of the learned boundary (default: the library's 0.02)")
    ap.add_argument("--config", choices=("c2", "c3", "c5", "sh"), default="c2",
                    help="c2: the 512x512x256 bench line (default); c3: the 1024x1024x512 reduced-storage "
                         "run against an FP64 run of the same scheme (one JSON line of its own); sh: the "
                         "8192x4096 SH (swpc_sh) step, FP32 and FP64 (one JSON line of its own); c5: the 20-cell "
                         "PML and the learned boundary of the c2 workload against an echo-free reference on an "
                         "extended grid (one JSON line of its own)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the solver holds the whole domain on one MI355X "
              "(multi-GPU sharding is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from openswpc_amd import problem
    from openswpc_amd.solver import Swpc3D

    if args.config == "c3":
        return run_c3(args)
    if args.config == "sh":
        return run_sh(args)
    if args.config == "c5":
        return run_c5(args)
    nt = args.warmup + args.steps
    t0 = time.time()
    # the source table covers both the timed run and the misfit window
    case = problem.case3d(args.nx, args.ny, args.nz, 0.25, 0.01, max(nt, args.misfit_steps), na=20, seed=args.seed,
                          rmed=(0.05, 2.0, 0.5), src_depth=10.0, nst_side=8, rmed_tile=(128, 128, 128))
    g = case.grid
    log(f"setup {time.time() - t0:.1f} s: grid {g.nx}x{g.ny}x{g.nz}, vmax {case.medium.vmax:.2f} km/s, "
        f"courant {problem.stability3(g, case.medium.vmax):.3f}, {len(case.stations)} receivers")

    n_int = g.n_interior_cells()
    n_pml = g.n_pml_cells()
    ncell = g.nx * g.ny * g.nz
    assert n_int + n_pml == ncell

    import ctypes as C
    from openswpc_amd._lib import lib
    gbs = C.c_double(0)
    copy_bw = None
    if lib().swpc_membw(1 << 30, 20, C.byref(gbs)) == 0:
        copy_bw = gbs.value
        log(f"measured copy bandwidth (1 GiB float4 copy): {copy_bw:.0f} GB/s")

    plain = timed_run(case, "fp32", args.warmup, args.steps)
    wall, ms, cnt = plain["wall"], plain["ms"], plain["cnt"]
    ms_step = wall / args.steps * 1e3
    cups = ncell * args.steps / wall
    # dominant kernel: the interior stress update (class 0: one launch of
    # the one-pass normal + shear kernel per step)
    avg_stress = ms[0] / max(cnt[0], 1) * 1e-3
    alg_stress = B_INT_STRESS * n_int
    achieved = alg_stress / avg_stress
    need_stress = B_INT_STRESS_NEED - 4 * plane_constant_arrays(case)
    step_alg = (B_INT_STRESS + B_INT_VEL) * n_int + (B_PML_STRESS + B_PML_VEL) * n_pml
    kern = {name: {"avg_us": ms[q] / max(cnt[q], 1) * 1e3, "launches": int(cnt[q])}
            for q, name in enumerate(["stress_int", "stress_pml", "vel_int", "vel_pml", "other"])}
    log(f"plain: {args.steps} steps in {wall:.3f} s, {ms_step:.3f} ms/step, {cups / 1e9:.2f} G cell-updates/s, "
        f"step {step_alg * args.steps / wall / 1e12:.2f} TB/s algorithmic; kernels {kern}")

    chosen = None
    if not args.no_chosen:
        from openswpc_amd import lab
        red = timed_run(case, CHOSEN, args.warmup, args.steps)
        ms_red = red["wall"] / args.steps * 1e3
        weights = lab.load_weights()[0] if lab.WEIGHTS.exists() else None
        icase = lab.interior_case(case) if weights is not None else None
        labr = timed_run(icase, CHOSEN, args.warmup, args.steps, lab=weights) if weights is not None else None
        lab32 = timed_run(icase, "fp32", args.warmup, args.steps, lab=weights) if weights is not None else None
        if args.no_misfit:
            # profiling runs: timings only, no accuracy and no choice
            mis = {"steps": 0, "fp32": None, "r16": None, "lab_r16": None, "lab_fp64": None,
                   "lab_r16_same_scheme": None, "lab_fp32_same_scheme": None}
        else:
            mis = misfit(case, args.misfit_steps, icase, weights)
        reduced = {"form": "reduced storage (fp32_r16): FP32 arithmetic, V, S, medium and ADE arrays; memory "
                           "variables in FP16 with a power-of-two scale; 20-cell ADE-PML",
                   "ms_per_step": ms_red, "value": ncell * args.steps / red["wall"],
                   "speedup_vs_plain": ms_step / ms_red, "misfit_vs_fp64": mis["r16"]}
        log(f"reduced storage: {ms_red:.3f} ms/step, speedup {ms_step / ms_red:.2f}x, misfit {mis['r16']}")
        chosen = {"misfit_steps": mis["steps"], "plain_misfit_vs_fp64": mis["fp32"], "reduced_storage": reduced,
                  "selection": f"the fastest form whose receiver traces are within {MISFIT_BAR:g} (relative L2) of the "
                               f"FP64 run of the plain scheme (20-cell PML) over the first {mis['steps']} steps"}
        forms = [("reduced_storage", reduced)]
        if labr is not None:
            gi = icase.grid
            ms_lab = labr["wall"] / args.steps * 1e3
            ms_lab32 = lab32["wall"] / args.steps * 1e3
            learned = {
                "form": f"learned absorbing boundary (per-face MLP ghost fill, BF16 weights, FP32 arithmetic) on the "
                        f"{gi.nx}x{gi.ny}x{gi.nz} interior of the plain grid (same physical region, receivers and "
                        f"source, no absorber cells) + reduced storage (fp32_r16)",
                "ms_per_step": ms_lab, "cells": gi.nx * gi.ny * gi.nz,
                "value": gi.nx * gi.ny * gi.nz * args.steps / labr["wall"],
                "value_note": "cell-updates/s over the cells this form updates (the 472x472x236 grid)",
                "physical_region_equivalent_cell_updates_per_s": ncell * args.steps / labr["wall"],
                "speedup_vs_plain": ms_step / ms_lab,
                "misfit_vs_fp64": mis["lab_r16"],
                "misfit_vs_fp64_note": "FP64 run of the plain scheme (20-cell PML): the two boundaries' difference "
                                       "once the reflections reach the receivers",
                "misfit_vs_fp64_same_scheme": mis["lab_r16_same_scheme"],
                "fp64_learned_vs_fp64_plain_pml20": mis["lab_fp64"],
                "lab_fp32_ms_per_step": ms_lab32, "lab_fp32_misfit_vs_fp64_same_scheme": mis["lab_fp32_same_scheme"],
                "lab_fill_us_per_step": (labr["ms"][1] / max(labr["cnt"][1], 1) + labr["ms"][3] / max(labr["cnt"][3], 1)) * 1e3,
            }
            forms.append(("learned_boundary", learned))
            chosen["learned_boundary"] = learned
            log(f"learned boundary + r16: {ms_lab:.3f} ms/step, speedup {ms_step / ms_lab:.2f}x, "
                f"misfit {mis['lab_r16']} vs the FP64 20-cell-PML run, {mis['lab_r16_same_scheme']} vs its "
                f"own FP64 run (learned boundary fp32: {ms_lab32:.3f} ms/step)")
        ok = [(f["ms_per_step"], k, f) for k, f in forms
              if f["misfit_vs_fp64"] is not None and f["misfit_vs_fp64"] <= MISFIT_BAR]
        if ok:
            _, key, best = min(ok)
            chosen.update({"chosen": key, "form": best["form"], "ms_per_step": best["ms_per_step"],
                           "value": best["value"], "speedup_vs_plain": best["speedup_vs_plain"],
                           "misfit_vs_fp64": best["misfit_vs_fp64"]})
            log(f"chosen: {key}, {best['ms_per_step']:.3f} ms/step, {best['speedup_vs_plain']:.2f}x")
        else:
            chosen["chosen"] = None

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(case, args.cpu_steps)
    c1 = None if args.no_c1 else run_c1(args.seed)

    out = {
        "metric": "cell-updates/s",
        "value": cups,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"swpc_3d plain viscoelastic {g.nx}x{g.ny}x{g.nz} dx=0.25km dt=10ms na=20 NM=3, "
                        f"layered + 5% von Karman medium, 1 double couple, {len(case.stations)} receivers",
            "nx": g.nx, "ny": g.ny, "nz": g.nz, "interior_cells": n_int, "pml_cells": n_pml,
            "parallelism": "single GPU, whole domain",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "interior stress update (k3x_stress, one launch per step)",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": traffic_from_profile(),
            "traffic_source": "FETCH_SIZE x 2 + WRITE_SIZE per launch from the rocprofv3 counter runs committed in "
                              "profiles/stress_int_512x512x256_traffic.json (a counter run of its own, not this one)",
            "alg_bytes_per_launch": alg_stress,
            "alg_bytes_per_cell": B_INT_STRESS,
            "kernel_streams_bytes_per_cell": need_stress,
            "kernel_streams_GBs": need_stress * n_int / avg_stress / 1e9,
            "avg_launch_us": avg_stress * 1e6,
            "measured_copy_GBs": copy_bw,
        },
        "step_roofline": {
            "alg_bytes_per_step": step_alg,
            "achieved_GBs": step_alg * args.steps / wall / 1e9,
            "frac": step_alg * args.steps / wall / HBM_PEAK,
        },
        "kernels": kern,
        "chosen": chosen,
        "cpu_baseline": cpu,
        "c1": c1,
    }
    print(json.dumps(out), flush=True)


def plane_constant_arrays(case):
    """Mean number, over the interior planes, of the stress update's
    medium arrays (lam, taup, taus) that the library reads from its
    per-plane table: a plane qualifies when every cell i, j in [1, n] holds
    the same value, bit for bit (swpc3d_set_medium).  Each such array is
    4 bytes per cell the kernel does not stream."""
    g, m = case.grid, case.medium
    tot = 0.0
    for a in (m.lam, m.taup, m.taus):
        u = np.ascontiguousarray(a[3:3 + g.ny, 3:3 + g.nx, :], np.float32).view(np.uint32)
        flags = [bool((u[:, :, k + 2] == u[0, 0, k + 2]).all()) for k in range(1, g.nz - g.na + 1)]
        tot += float(np.mean(flags))
    return tot


def timed_run(case, precision, warmup, steps, lab=None):
    """Build the context, run warmup steps, then time `steps` steps with the
    per-kernel HIP events enabled; inputs are resident in HBM throughout.
    lab: learned-boundary nets for a grid without PML."""
    from openswpc_amd.solver import Swpc3D
    g = case.grid
    nt = warmup + steps
    sim = Swpc3D(g, case.medium, case.pml, case.glut, case.stations, case.ntdec_w, nt, precision=precision)
    if lab is not None:
        sim.set_learned_boundary(lab)
    log(f"{precision}{' + learned boundary' if lab is not None else ''}: device memory {sim.memory() / 2**30:.2f} GiB")
    sim.run(1, warmup)
    sim.sync()
    sim.profile(True)
    t1 = time.perf_counter()
    sim.run(warmup + 1, steps)
    sim.sync()
    wall = time.perf_counter() - t1
    ms, cnt = sim.profile_read()
    sim.profile(False)
    vel, _ = sim.receivers()
    sim.close()
    if not np.all(np.isfinite(vel)):
        raise RuntimeError(f"{precision}: non-finite receiver output")
    return {"wall": wall, "ms": ms, "cnt": cnt}


def misfit(case, steps, icase=None, weights=None):
    """Receiver-trace relative L2 misfit against the FP64 run of the plain
    scheme (20-cell PML) over the first `steps` steps of the benchmark case:
    plain FP32, the reduced form, and the learned-boundary forms on the
    interior grid (same receivers at the same physical positions)."""
    from openswpc_amd.solver import Swpc3D
    runs = [("fp64", case, "fp64", None), ("fp32", case, "fp32", None), ("r16", case, CHOSEN, None)]
    if weights is not None:
        runs += [("lab_fp64", icase, "fp64", weights), ("lab_r16", icase, CHOSEN, weights),
                 ("lab_fp32", icase, "fp32", weights)]
    tr = {}
    for key, c, prec, w in runs:
        sim = Swpc3D(c.grid, c.medium, c.pml, c.glut, c.stations, c.ntdec_w, steps, precision=prec)
        if w is not None:
            sim.set_learned_boundary(w)
        sim.run(1, steps)
        tr[key] = sim.receivers()[0].astype(np.float64)
        sim.close()
    ref = tr["fp64"]
    out = {"steps": steps}
    for key in tr:
        if key != "fp64":
            out[key] = rel_misfit(tr[key], ref, f"{key} vs fp64 over {steps} steps")
    if "lab_fp64" in tr:
        # the learned-boundary forms against the FP64 run of the same scheme
        r = tr["lab_fp64"]
        out["lab_r16_same_scheme"] = rel_misfit(tr["lab_r16"], r, "lab_r16 vs lab_fp64")
        out["lab_fp32_same_scheme"] = rel_misfit(tr["lab_fp32"], r, "lab_fp32 vs lab_fp64")
    return out


def rel_misfit(a, ref, what):
    """Relative L2 misfit ||a - ref|| / ||ref||; an empty or non-finite
    reference is an error, never a NaN in the output line."""
    n = float(np.linalg.norm(ref))
    if not np.isfinite(n) or n == 0.0:
        raise RuntimeError(f"misfit {what}: the reference traces have norm {n} (no signal reached the "
                           f"receivers in the window, or the run diverged)")
    e = float(np.linalg.norm(np.asarray(a, np.float64) - ref)) / n
    if not np.isfinite(e):
        raise RuntimeError(f"misfit {what}: non-finite ({e})")
    return e


# Algorithmic HBM bytes per P-SV cell and step, FP32 (interior 38 words,
# PML 34 words: src/swpc_psv/m_kernel.f90 and m_absorb_p.f90 under the
# two-phase order).
B_PSV_INT = 38 * 4
B_PSV_PML = 34 * 4


def run_c1(seed=1, nt=4000):
    """The P-SV configuration run in full: 800 x 400 cells, dx = dz = 0.1 km,
    dt = 5 ms, 4000 steps, 20-cell PML, two flat layers (5.8/3.36/2.7 over
    6.5/3.75/2.9 below 15 km, Qp 600 / Qs 300, 3 Zener bodies), a double
    couple with a 1 s Kupper rise at 10 km depth (mechanism from the seed),
    16 surface receivers 2 km apart.  The GPU runs FP32 (timed) and FP64;
    the FP32 oracle runs the same 4000 steps on the host cores (timed: the
    CPU baseline of this configuration).  Traces are compared FP32 GPU vs
    FP32 oracle and vs the FP64 GPU run."""
    from openswpc_amd import problem
    from openswpc_amd.solver import SwpcPSV
    case = problem.case2d(800, 400, 0.1, 0.005, nt, na=20, seed=seed, src_depth=10.0, nst=16, st_spacing=2.0)
    g = case.grid
    ncell = g.nx * g.nz
    n_int = (g.nx - 2 * g.na) * (g.nz - g.na)
    n_pml = ncell - n_int
    out = {"workload": "swpc_psv 800x400 dx=dz=0.1km dt=5ms nt=4000 na=20 NM=3, two flat layers, "
                       "1 double couple at 10 km, 16 surface receivers (run in full)",
           "steps": nt}
    tr = {}
    for prec in ("fp32", "fp64"):
        with SwpcPSV(g, case.medium, case.pml, case.glut, case.stations, case.ntdec_w, nt, precision=prec) as sim:
            sim.run(1, 20)
            sim.sync()
            t1 = time.perf_counter()
            sim.run(21, nt - 20)
            sim.sync()
            wall = time.perf_counter() - t1
            tr[prec] = sim.receivers()[0].astype(np.float64)
        if prec == "fp32":
            step = wall / (nt - 20)
            alg = B_PSV_INT * n_int + B_PSV_PML * n_pml
            out.update({"ms_per_step": step * 1e3, "value": ncell / step, "unit": "cell-updates/s",
                        "alg_bytes_per_step": alg, "alg_GBs": alg / step / 1e9,
                        "hbm_frac_note": "the ~47 MB state fits in the 256 MB Infinity Cache: "
                                         "reported, not an HBM-bound figure",
                        "alg_frac_of_hbm_peak": alg / step / HBM_PEAK})
        else:
            out["fp64_ms_per_step"] = wall / (nt - 20) * 1e3
    cpu = None
    try:
        from oracle.oracle import Oracle2D, olib
        ref = Oracle2D(case, 32)
        nth = olib(32).oracle_num_threads()
        t1 = time.perf_counter()
        ref.run(1, nt)
        cwall = time.perf_counter() - t1
        cpu = {"value": ncell * nt / cwall, "unit": "cell-updates/s", "cores": int(nth), "kind": "port",
               "sample": "the whole 4000-step run, FP32 oracle"}
        out["misfit_fp32_vs_oracle_fp32"] = rel_misfit(tr["fp32"], ref.vel.astype(np.float64),
                                                       "c1 fp32 vs oracle")
    except ImportError as e:  # pragma: no cover
        log(f"c1: oracle unavailable ({e})")
    out["misfit_fp32_vs_fp64"] = rel_misfit(tr["fp32"], tr["fp64"], "c1 fp32 vs fp64")
    out["cpu_baseline"] = cpu
    log(f"c1: {out}")
    return out


def run_c3(args, nt=4000, plain_steps=300, chunk=250):
    """1024 x 1024 x 512 cells, dx = 0.125 km, dt = 5 ms, 4000 steps: the
    reduced-storage form (fp32_r16) timed over the whole run, its receiver
    traces against an FP64 run of the same scheme (20-cell PML), and the
    plain FP32 form timed over a short window for its step time."""
    from openswpc_amd import problem
    from openswpc_amd.solver import Swpc3D
    t0 = time.time()
    case = problem.case3d(1024, 1024, 512, 0.125, 0.005, nt, na=20, seed=args.seed, rmed=(0.03, 1.0, 0.5),
                          src_depth=10.0, nst_side=8, rmed_tile=(128, 128, 128), log=log, device_medium=True)
    g = case.grid
    ncell = g.nx * g.ny * g.nz
    host_setup = time.time() - t0
    log(f"c3 host setup (velocity model) {host_setup:.0f} s; medium__setup runs on the GPU")
    out = {"config": {"workload": "swpc_3d 1024x1024x512 dx=0.125km dt=5ms na=20 NM=3, layered + 3% von Karman "
                                  "(a=1 km), 1 double couple, 64 receivers", "nt": nt},
           "setup": {"host_velocity_model_s": host_setup}}
    traces = {}
    for prec in (CHOSEN, "fp64"):
        t1 = time.perf_counter()
        sim = problem.open_case3d(case, prec)
        sim.sync()
        out["setup"][f"device_setup_{prec}_s"] = time.perf_counter() - t1
        log(f"c3 {prec}: context with the medium set up on the GPU in {time.perf_counter() - t1:.1f} s, "
            f"vmax {case.medium.vmax:.2f} km/s, courant {problem.stability3(g, case.medium.vmax):.3f}, "
            f"device memory {sim.memory() / 2**30:.1f} GiB")
        sim.sync()
        t1 = time.perf_counter()
        for it in range(1, nt + 1, chunk):
            sim.run(it, min(chunk, nt - it + 1))
            sim.sync()
            log(f"c3 {prec}: {min(it + chunk - 1, nt)}/{nt} steps, {time.perf_counter() - t1:.1f} s")
        wall = time.perf_counter() - t1
        traces[prec] = sim.receivers()[0].astype(np.float64)
        out[prec] = {"ms_per_step": wall / nt * 1e3, "cell_updates_per_s": ncell * nt / wall,
                     "device_GiB": sim.memory() / 2**30}
        sim.close()
    sim = problem.open_case3d(case, "fp32")
    sim.run(1, 20)
    sim.sync()
    t1 = time.perf_counter()
    sim.run(21, plain_steps)
    sim.sync()
    wall = time.perf_counter() - t1
    out["fp32"] = {"ms_per_step": wall / plain_steps * 1e3, "cell_updates_per_s": ncell * plain_steps / wall,
                   "device_GiB": sim.memory() / 2**30, "timed_steps": plain_steps}
    sim.close()
    ref = traces["fp64"]
    out["misfit_r16_vs_fp64"] = float(np.linalg.norm(traces[CHOSEN] - ref) / np.linalg.norm(ref))
    out["speedup_r16_vs_fp32"] = out["fp32"]["ms_per_step"] / out[CHOSEN]["ms_per_step"]
    log(f"c3: {out}")
    print(json.dumps(out), flush=True)


def run_c5(args, nt=2000, ext=256, ext_z=200, chunk=250, early=700):
    """The boundaries of the 512 x 512 x 256 workload against an echo-free
    reference over nt steps: the same medium, source and receivers on a
    grid extended by ext cells on each side in x and y and ext_z cells at
    the bottom (1024 x 1024 x 456 by default), far enough that no wave
    reflected by its own absorber reaches a receiver inside the window
    (path length / vmax > nt dt).  Receiver traces of the 20-cell PML
    (FP32, FP64) and of the learned boundary on the 472 x 472 x 236 interior
    (reduced storage, FP32, FP64) against it: relative L2 misfit over the
    whole window, over the first `early` steps (before the first boundary
    reflection reaches the nearest receiver, which checks that the grids
    agree) and over the rest (the reflections)."""
    from openswpc_amd import lab, problem
    from openswpc_amd.solver import Swpc3D
    if not lab.WEIGHTS.exists():
        raise RuntimeError("c5: no trained nets in the package")
    weights = lab.load_weights()[0]
    sp = (512 - 2 * 20) * 0.25 / 9   # the receiver spacing of the 512 grid's default 8 x 8 layout
    kw = dict(na=20, seed=args.seed, rmed=(0.05, 2.0, 0.5), src_depth=10.0, nst_side=8, st_spacing=sp,
              rmed_tile=(128, 128, 128))
    t0 = time.time()
    case = problem.case3d(512, 512, 256, 0.25, 0.01, nt, **kw)
    icase = lab.interior_case(case)
    big = problem.case3d(512 + 2 * ext, 512 + 2 * ext, 256 + ext_z, 0.25, 0.01, nt, device_medium=True, **kw)
    gb = big.grid
    vmax = case.medium.vmax
    # shortest path source -> own absorber -> receiver, in km
    xb = (gb.nx // 2 - gb.na) * gb.dx
    zb = gb.zbeg + (gb.nz - gb.na) * gb.dz
    rmax = max(abs(s.x) for s in case.station_list)
    path = min(2 * xb - rmax, 2 * zb - 10.0)
    log(f"c5 setup {time.time() - t0:.0f} s; reference grid {gb.nx}x{gb.ny}x{gb.nz}: its own reflections need "
        f"{path / vmax:.1f} s > window {nt * 0.01:.1f} s")
    if path / vmax <= nt * 0.01:
        raise RuntimeError("c5: the reference grid is too small for the window")

    def run(c, prec, w=None, raw=False):
        t1 = time.perf_counter()
        sim = problem.open_case3d(c, prec) if raw else Swpc3D(c.grid, c.medium, c.pml, c.glut, c.stations,
                                                              c.ntdec_w, nt, precision=prec)
        if w is not None:
            sim.set_learned_boundary(w)
            if args.lab_sponge is not None:
                sim.set_learned_boundary_sponge(args.lab_sponge)
        for it in range(1, nt + 1, chunk):
            sim.run(it, min(chunk, nt - it + 1))
            sim.sync()
            log(f"c5 {c.grid.nx}x{c.grid.ny}x{c.grid.nz} {prec}{' + learned boundary' if w is not None else ''}: "
                f"{min(it + chunk - 1, nt)}/{nt} steps, {time.perf_counter() - t1:.1f} s")
        v = sim.receivers()[0].astype(np.float64)
        sim.close()
        if not np.all(np.isfinite(v)):
            raise RuntimeError("c5: non-finite traces")
        return v

    truth = run(big, "fp32", raw=True)
    names = [s.name for s in case.station_list]
    if [s.name for s in big.station_list] != names or [s.name for s in icase.station_list] != names:
        raise RuntimeError("c5: the receiver lists differ")
    out = {"config": {"workload": f"swpc_3d 512x512x256 workload's boundaries vs an echo-free {gb.nx}x{gb.ny}x{gb.nz} "
                                  f"reference, {nt} steps", "nt": nt, "early_steps": early,
                      "reference_grid": [gb.nx, gb.ny, gb.nz],
                      "reference_own_reflection_s": path / vmax,
                      "lab_sponge": 0.02 if args.lab_sponge is None else args.lab_sponge},
           "runs": {}}
    forms = [("pml20_fp32", case, "fp32", None), ("pml20_fp64", case, "fp64", None),
             ("learned_r16", icase, CHOSEN, weights), ("learned_fp32", icase, "fp32", weights),
             ("learned_fp64", icase, "fp64", weights)]
    for key, c, prec, w in forms:
        v = run(c, prec, w)
        r = {"misfit": rel_misfit(v, truth, f"{key} vs reference"),
             "misfit_early": rel_misfit(v[:early], truth[:early], f"{key} early"),
             "misfit_late": rel_misfit(v[early:], truth[early:], f"{key} late"),
             "residual_energy_late_over_total": float(np.sum((v[early:] - truth[early:]) ** 2) / np.sum(truth ** 2))}
        per = np.linalg.norm((v - truth).reshape(nt, -1, 3), axis=(0, 2)) / np.linalg.norm(truth.reshape(nt, -1, 3),
                                                                                            axis=(0, 2))
        r["worst_receiver"] = {"name": names[int(np.argmax(per))], "misfit": float(per.max())}
        out["runs"][key] = r
        log(f"c5 {key}: {r}")
    print(json.dumps(out), flush=True)


# Algorithmic HBM bytes per interior SH cell and step, FP32 words
# (src/swpc_sh/m_kernel.f90): stress reads Vy, mu, taus, 6 memory variables,
# Syz, Sxy and writes the 6 and the 2 stresses (19 words); velocity reads
# Syz, Sxy, rho, Vy and writes Vy (5 words).
SH_CELL_WORDS = 19 + 5


def run_sh(args, nx=8192, nz=4096, chunk=100):
    """The SH core's step time on an 8192 x 4096 grid (dx = 0.1 km,
    dt = 5 ms, na = 20, NM = 3, layered medium, one double couple, 16
    receivers) in FP32 and FP64, steps replayed from captured graphs."""
    from openswpc_amd import problem
    from openswpc_amd.solver import SwpcSH
    log = lambda m: print(m, file=sys.stderr, flush=True)   # noqa: E731
    steps = min(args.steps, 1000)
    nt = args.warmup + steps
    case = problem.case_sh(nx, nz, 0.1, 0.005, nt, na=20, seed=args.seed, src_depth=10.0, nst=16, zbeg=-1.0)
    interior = (nx - 2 * case.grid.na) * (nz - case.grid.na)
    out = {"config": {"workload": f"swpc_sh {nx}x{nz} dx=0.1km dt=5ms na=20 NM=3, layered, 1 double couple, "
                                  "16 receivers", "timed_steps": steps}}
    for prec in ("fp32", "fp64"):
        with SwpcSH(case.grid, case.medium, case.pml, case.glut, case.stations, case.ntdec_w, nt,
                    precision=prec) as sim:
            sim.run(1, args.warmup)
            sim.sync()
            t1 = time.perf_counter()
            for it in range(args.warmup + 1, nt + 1, chunk):
                sim.run(it, min(chunk, nt + 1 - it))
            sim.sync()
            dt = (time.perf_counter() - t1) / steps
            word = 4 if prec == "fp32" else 8
            # FP64 fields; the medium stays FP32 (mu, taus, rho: 3 words)
            nbytes = SH_CELL_WORDS * word - (3 * 4 if prec == "fp64" else 0)
            out[prec] = {"ms_per_step": dt * 1e3, "cell_updates_per_s": nx * nz / dt,
                         "interior_bytes_per_cell": nbytes,
                         "interior_algorithmic_GBs": interior * nbytes / dt / 1e9,
                         "hbm_peak_frac": interior * nbytes / dt / 8e12}
            log(f"sh {prec}: {out[prec]}")
    print(json.dumps(out), flush=True)


def traffic_from_profile():
    """HBM bytes per interior-stress update from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary, if present."""
    p = ROOT / "profiles" / "stress_int_512x512x256_traffic.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text()).get("bytes_per_step_update")
    except Exception:
        return None


def cpu_baseline(case, steps):
    """Oracle FP32 (C/OpenMP restatement of the reference loops) on the same
    grid for a bounded number of steps, all host cores."""
    try:
        from oracle.oracle import Oracle3D, olib
    except Exception as e:  # pragma: no cover
        log(f"cpu baseline unavailable: {e}")
        return None
    ncores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    ref = Oracle3D(case, 32)
    nth = olib(32).oracle_num_threads()
    ref.run(1, 1)
    t = time.perf_counter()
    ref.run(2, steps)
    dt = time.perf_counter() - t
    g = case.grid
    v = g.nx * g.ny * g.nz * steps / dt
    log(f"cpu baseline: {steps} steps in {dt:.2f} s on {nth} threads ({ncores} cores visible): {v / 1e6:.1f} M cell-updates/s")
    # the same steps on the GPU: every field against the oracle's
    from openswpc_amd.solver import Swpc3D
    per = {}
    with Swpc3D(g, case.medium, case.pml, case.glut, case.stations, case.ntdec_w, 1 + steps) as sim:
        sim.run(1, 1 + steps)
        for f in ("Vx", "Vy", "Vz", "Sxx", "Syy", "Szz", "Syz", "Sxz", "Sxy", "Rxx1", "Rxy3"):
            per[f] = rel_misfit(sim.field(f), ref.field(f), f"field {f}")
    log(f"GPU vs oracle after {1 + steps} steps: max field rel-L2 {max(per.values()):.2e}")
    return {"value": v, "unit": "cell-updates/s", "cores": int(nth), "kind": "port",
            "sample": f"{steps} full steps (after 1 untimed) of the same {g.nx}x{g.ny}x{g.nz} case, FP32 oracle",
            "gpu_vs_oracle": {"steps": 1 + steps, "max_field_rel_l2": max(per.values()), "fields_rel_l2": per}}


if __name__ == "__main__":
    main()
