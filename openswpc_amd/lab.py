"""Learned absorbing boundary: per-face ghost-fill MLPs on the GPU.

No counterpart in the reference: this is the thin absorbing boundary that
takes the place of the ADE-PML (``absorb_p__update_stress/vel``,
src/swpc_3d/m_absorb_p.f90:126-311, 313-505) on a grid created with
``na = 0``.  The interior update covers every cell; before each half-step
the ghost cells beyond the x-, x+, y-, y+ and bottom faces are written by a
small MLP per face and half-step from the six outermost cells of each
column and the net's own previous ghost values (the definition is stated
in openswpc_amd/csrc/swpc_lab.h and include/swpc_hip.h).

Training data come from the solver itself: reference runs on a grid 110
cells wider at each absorbing face, with reflecting edges far enough away
that no echo of them re-enters the inner box within the run (an
echo-free stand-in for the unbounded medium); the inner box stands for a
grid with a learned boundary, and the reference run's fields in that
box's ghost cells are the targets.  ``Trainer`` holds the
samples on the GPU and fits the nets with the library's own forward,
backward and Adam kernels.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import fdtool, medium as md, pml, problem, rmedia
from ._lib import check, lib, ptr
from .source import PointSource, Station, station_cells3, stress_glut3

FACES = ("x-", "x+", "y-", "y+", "z+")
NNET = 10                       # net = 2*face + phase
NOUT = (4, 5, 5, 4, 4, 5, 5, 4, 5, 4)
NIN, NH, NO = 61, 32, 5
NSC = 59                        # features scaled by 1/N
REC = 72                        # floats per training sample
REF_NA = 0                      # reference runs: no absorber, the edges are far away
OFFSET = 110                    # inner box offset inside the reference grid (cells)
SPONGE = 0.02                   # default strip sponge strength of the time step
WEIGHTS = Path(__file__).resolve().parent / "data" / "lab_weights.npz"


def offsets():
    """Parameter offsets of one net: W1[32][61] b1 W2[32][32] b2 W3[5][32] b3 Ws[5][61]."""
    o = {"W1": 0}
    o["b1"] = o["W1"] + NH * NIN
    o["W2"] = o["b1"] + NH
    o["b2"] = o["W2"] + NH * NH
    o["W3"] = o["b2"] + NH
    o["b3"] = o["W3"] + NO * NH
    o["Ws"] = o["b3"] + NO
    o["n"] = o["Ws"] + NO * NIN
    return o


NP = offsets()["n"]


def unpack(p):
    """Views (W1, b1, W2, b2, W3, b3, Ws) of one net's flat parameters."""
    o = offsets()
    return (p[o["W1"]:o["b1"]].reshape(NH, NIN), p[o["b1"]:o["W2"]], p[o["W2"]:o["b2"]].reshape(NH, NH),
            p[o["b2"]:o["W3"]], p[o["W3"]:o["b3"]].reshape(NO, NH), p[o["b3"]:o["Ws"]],
            p[o["Ws"]:o["n"]].reshape(NO, NIN))


def init_params(seed=0):
    """Ten nets: hidden layers ~ N(0, 1/fan_in), output and skip layers
    zero, so an untrained boundary writes zero ghosts."""
    rng = np.random.default_rng(seed)
    P = np.zeros((NNET, NP), np.float32)
    for n in range(NNET):
        W1, b1, W2, b2, W3, b3, Ws = unpack(P[n])
        W1[...] = rng.standard_normal(W1.shape) / np.sqrt(NIN)
        W2[...] = rng.standard_normal(W2.shape) / np.sqrt(NH)
    return P


def bf16_round(a):
    """FP32 -> nearest-even BF16, returned as FP32 (what the time step uses)."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(np.shape(a))


def save_weights(path, params, **meta):
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez(path, params=np.asarray(params, np.float32),
             **{k: np.asarray(v) for k, v in meta.items()})


def load_weights(path=WEIGHTS):
    with np.load(path, allow_pickle=False) as z:
        P = np.array(z["params"], np.float32)
        meta = {k: z[k] for k in z.files if k != "params"}
    if P.shape != (NNET, NP):
        raise ValueError(f"{path}: parameters of shape {P.shape}, expected {(NNET, NP)}")
    return P, meta


class Trainer:
    """Sample store and optimiser on the GPU (swpc_lab_trainer)."""

    def __init__(self, cap):
        h = C.c_void_p()
        check(lib().swpc_lab_trainer_create(int(cap), C.byref(h)), "swpc_lab_trainer_create")
        self.h = h
        self.cap = int(cap)

    def close(self):
        if getattr(self, "h", None):
            lib().swpc_lab_trainer_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def begin_run(self, sim, oi, oj, nxs, nys, nzs, quota=None):
        q = None if quota is None else np.ascontiguousarray(np.broadcast_to(quota, (NNET,)), np.int32)
        check(lib().swpc_lab_begin_run(self.h, sim.h, oi, oj, nxs, nys, nzs, ptr(q, C.c_int)), "swpc_lab_begin_run")

    def gather(self, sim, phase, keep=1.0, seed=0, thr=0.0):
        check(lib().swpc3d_lab_gather(sim.h, self.h, phase, float(keep), int(seed) & 0xFFFFFFFF, float(thr)),
              "swpc3d_lab_gather")

    def counts(self):
        c = np.zeros(NNET, np.int32)
        check(lib().swpc_lab_count(self.h, ptr(c, C.c_int)), "swpc_lab_count")
        return c

    def add(self, net, rec):
        rec = np.ascontiguousarray(rec, np.float32).reshape(-1, REC)
        check(lib().swpc_lab_add(self.h, net, len(rec), ptr(rec, C.c_float)), "swpc_lab_add")

    def samples(self, net, first=0, n=None):
        if n is None:
            n = int(self.counts()[net]) - first
        out = np.zeros((n, REC), np.float32)
        if n:
            check(lib().swpc_lab_samples(self.h, net, first, n, ptr(out, C.c_float)), "swpc_lab_samples")
        return out

    def train(self, net, params, steps, batch=8192, lr=3e-3, noise=0.1, seed=0):
        p = np.ascontiguousarray(params, np.float32).copy()
        loss = np.zeros(max(steps, 1), np.float32)
        check(lib().swpc_lab_train(self.h, net, ptr(p, C.c_float), steps, batch, float(lr), float(noise),
                                   int(seed) & 0xFFFFFFFF, ptr(loss, C.c_float)), "swpc_lab_train")
        return p, loss[:steps]

    def grad(self, params, nout, rec):
        p = np.ascontiguousarray(params, np.float32)
        rec = np.ascontiguousarray(rec, np.float32).reshape(-1, REC)
        g = np.zeros(NP, np.float32)
        loss = C.c_double(0)
        check(lib().swpc_lab_grad(self.h, ptr(p, C.c_float), nout, len(rec), ptr(rec, C.c_float), C.byref(loss),
                                  ptr(g, C.c_float)), "swpc_lab_grad")
        return loss.value, g


# ---------------------------------------------------------------------------
# cases: reference runs and the learned-boundary grid inside them
# ---------------------------------------------------------------------------
@dataclass
class LabCase:
    wide: problem.Case3D       # the reference run (training targets, truth of the evaluation)
    src: PointSource
    nxs: int
    nys: int
    nzs: int
    raw: tuple = None          # (rho, lam, mu, qp, qs) of the wide grid before medium__setup
    off: int = OFFSET          # inner box offset inside the wide grid


def wide_case(nxs, nys, nzs, dx, dt, nt, seed, zbeg=-0.75, eps=0.05, a=2.0, layers=problem.LHM_EXAMPLE,
              off=OFFSET, na=REF_NA, device_medium=False):
    """The reference run of one training or evaluation case: a grid of
    (nxs + 2*off) x (nys + 2*off) x (nzs + off) cells (absorber of na cells,
    none by default) holding the inner nxs x nys x nzs box at (off, off, 0);
    the layered medium of ``layers`` with a von Karman perturbation (eps, a
    km, kappa 0.5) and a double couple of random position, depth and
    mechanism inside the inner box, all drawn from the seed.  With na = 0
    the run stands for the unbounded medium while no echo of its edges can
    reach the inner box: 2*off cells of travel, i.e. nt*vmax*dt/dx < 2*off.
    device_medium=True leaves medium__setup to the GPU (the training runs
    need no host copy of the assembled medium; inner_case does)."""
    g = md.Grid3D(nxs + 2 * off, nys + 2 * off, nzs + off, dx, dx, dx, dt, na=na, zbeg=zbeg)
    xi = None
    if eps > 0:
        tile = tuple(min(128, n) for n in (g.nx, g.ny, g.nz))
        vol = rmedia.generate(tile, dx, a, eps, 0.5, seed=seed)
        xi = md.rmed_tile(vol, -2, g.nx + 3, -2, g.ny + 3, -2, g.nz + 3)
    rho, lam, mu, qp, qs = md.layered3(g, layers, xi=xi)
    if device_medium:
        med = md.RawMedium3D(rho, lam, mu, qp, qs)
    else:
        med = md.medium_setup3(g, rho, lam, mu, qp, qs)
    rng = np.random.default_rng(seed + 7919)
    mech = (float(rng.uniform(0, 360)), float(rng.uniform(0, 90)), float(rng.uniform(-180, 180)))
    hx, hy = (nxs / 2 - 8) * dx, (nys / 2 - 8) * dx
    x, y = float(rng.uniform(-hx, hx)), float(rng.uniform(-hy, hy))
    z = float(rng.uniform(max(1.0, zbeg + 8 * dx), zbeg + (nzs - 8) * dx))
    src = PointSource(x, y, z, 0.1, 1.0, 1e15, fdtool.sdr2moment(*mech))
    glut = stress_glut3(g, [src], nt)
    if na > 0:
        prof = pml.pml3(g, glut.fcut)
    else:
        prof = tuple(np.zeros((n, 4), np.float32) for n in (g.nx, g.nx, g.ny, g.ny, g.nz, g.nz))
    case = problem.Case3D(g, med, prof, glut, np.zeros((0, 3), np.int32), [], nt, 1, mech)
    return LabCase(case, src, nxs, nys, nzs, (rho, lam, mu, qp, qs), off)


def _stations(lc: LabCase, nst_side):
    dx = lc.wide.grid.dx
    sp = (lc.nxs - 2 * 8) * dx / (nst_side + 1)
    return [Station((a_ - (nst_side - 1) / 2) * sp + 0.3 * dx, (b_ - (nst_side - 1) / 2) * sp + 0.3 * dx, 0.0,
                    f"L{a_:02d}{b_:02d}", "fsb") for a_ in range(nst_side) for b_ in range(nst_side)]


def inner_case(lc: LabCase, na=0, nst_side=0, grow=0):
    """The inner box of a wide run (grown by ``grow`` cells at each
    absorbing face) as a grid of its own, with the same source and time
    step and nst_side^2 receivers on the free surface at fixed physical
    positions.  With a learned boundary (na = 0) the medium is sliced from
    the wide grid's assembled medium, so that its interior update equals
    the wide run's; with a PML of na cells it is assembled by
    medium__setup from the sliced raw model, which homogenises it across
    the absorber as the reference does (m_medium.f90:124-190)."""
    wg, wm = lc.wide.grid, lc.wide.medium
    nx, ny, nz = lc.nxs + 2 * grow, lc.nys + 2 * grow, lc.nzs + grow
    o = lc.off - grow
    g = md.Grid3D(nx, ny, nz, wg.dx, wg.dy, wg.dz, wg.dt, na=na, zbeg=wg.zbeg)
    sj, si, sk = slice(o, o + ny + 6), slice(o, o + nx + 6), slice(0, nz + 6)
    if na > 0 and lc.raw is not None:
        med = md.medium_setup3(g, *[np.ascontiguousarray(a[sj, si, sk]) for a in lc.raw])
    else:
        kw = {n: np.ascontiguousarray(getattr(wm, n)[sj, si, sk])
              for n in ("rho", "lam", "mu", "taup", "taus", "bx", "by", "bz", "muyz", "muxz", "muxy")}
        kw.update({n: np.ascontiguousarray(getattr(wm, n)[sj, si])
                   for n in ("kfs", "kob", "kfs_top", "kfs_bot", "kob_top", "kob_bot")})
        med = md.Medium3D(**kw, ts=wm.ts, zeta=wm.zeta, vmin=wm.vmin, vmax=wm.vmax)
    glut = stress_glut3(g, [lc.src], lc.wide.nt)
    if na > 0:
        prof = pml.pml3(g, glut.fcut)
    else:
        prof = tuple(np.zeros((n, 4), np.float32) for n in (g.nx, g.nx, g.ny, g.ny, g.nz, g.nz))
    cells, kept = np.zeros((0, 3), np.int32), []
    if nst_side:
        cells, kept = station_cells3(g, med.kfs, med.kob, _stations(lc, nst_side))
    return problem.Case3D(g, med, prof, glut, cells, kept, lc.wide.nt, 1, lc.wide.mech)


def gather_run(trainer, lc: LabCase, quota=None, keep=1.0, thr=1e-3, seed=0, precision="fp32"):
    """One reference run on the GPU with the inner box's ghost targets and
    features gathered at every half-step."""
    from .solver import Swpc3D
    c = lc.wide
    with Swpc3D(c.grid, c.medium, c.pml, c.glut, None, 1, c.nt, precision=precision) as sim:
        trainer.begin_run(sim, lc.off, lc.off, lc.nxs, lc.nys, lc.nzs, quota)
        for it in range(1, c.nt + 1):
            s = (seed * 1_000_003 + it) & 0xFFFFFFFF
            trainer.gather(sim, 0, keep, s, thr)
            sim.kernel__update_stress()
            sim.source__stressglut(it)
            trainer.gather(sim, 1, keep, s ^ 0x5BD1E995, thr)
            sim.kernel__update_vel()
        sim.sync()


def random_layers(rng):
    """A random 1-D model for training variety: 3-5 layers with P speeds
    rising from 4.5-6 to 7.5-8.6 km/s, Vs = Vp/sqrt(3) within 6 %, density
    after Gardner (1.74 Vp^0.25), Qp 600, Qs 300."""
    n = int(rng.integers(3, 6))
    tops = np.concatenate([[0.0], np.sort(rng.uniform(1.0, 45.0, n - 1))])
    vps = np.sort(rng.uniform(4.5, 8.6, n))
    vps[0] = min(vps[0], 6.0)
    out = []
    for z, vp in zip(tops, vps):
        vs = vp / np.sqrt(3.0) * rng.uniform(0.94, 1.06)
        out.append(md.Layer(float(z), float(1.74 * vp ** 0.25), float(vp), float(vs), 600.0, 300.0))
    return out


def train_boundary(nruns=200, nxs=64, nys=64, nzs=48, dx=0.25, dt=0.01, nt=500, off=OFFSET, cap=2_000_000,
                   steps=6000, batch=8192, lr=3e-3, noise=0.1, keep=0.012, thr=1e-3, seed=1, log=print):
    """Generate nruns reference runs on the GPU (even runs: the example's
    layered model at a random depth offset; odd runs: a random layered
    model under a free surface; all with a 5 % von Karman perturbation and
    a random double couple), gather their ghost targets, and fit the ten
    nets.  Each run's inner nxs x nys x nzs box sits off cells inside its
    reflecting edges, which must exceed half the distance the fastest wave
    covers in nt steps.  Returns (params[10][NP], report)."""
    import time
    from concurrent.futures import ThreadPoolExecutor
    t0 = time.time()
    tr = Trainer(cap)
    quota = cap // nruns
    rng = np.random.default_rng(seed)
    specs = []
    for r in range(nruns):
        if r % 2 == 0:
            layers, zbeg = problem.LHM_EXAMPLE, float(rng.uniform(-0.75, 50.0))
        else:
            layers, zbeg = random_layers(rng), -0.75
        specs.append((int(seed) * 100_003 + r, layers, float(np.float32(round(zbeg / dx) * dx))))
    # the host builds the next cases (FFT random media, layered model)
    # while the GPU runs the current one; medium__setup runs on the GPU
    with ThreadPoolExecutor(max_workers=6) as pool:
        futs = [pool.submit(wide_case, nxs, nys, nzs, dx, dt, nt, s, zbeg, 0.05, 2.0, layers, off, REF_NA, True)
                for s, layers, zbeg in specs[:4]]
        for r in range(nruns):
            lc = futs[r].result()
            futs[r] = None
            if r + 4 < nruns:
                s, layers, zbeg = specs[r + 4]
                futs.append(pool.submit(wide_case, nxs, nys, nzs, dx, dt, nt, s, zbeg, 0.05, 2.0, layers, off, REF_NA, True))
            gather_run(tr, lc, quota, keep, thr, specs[r][0])
            del lc
            if log and (r % 20 == 19 or r == nruns - 1):
                log(f"lab: {r + 1}/{nruns} wide runs gathered, samples per net {tr.counts().tolist()}, "
                    f"{time.time() - t0:.0f} s")
    counts = tr.counts()
    P = init_params(seed)
    final = []
    for net in range(NNET):
        P[net], loss = tr.train(net, P[net], steps, batch, lr, noise, seed * 31 + net)
        final.append(float(loss[-200:].mean()))
        if log:
            log(f"lab: net {net} ({FACES[net // 2]} phase {net % 2}): {counts[net]} samples, "
                f"loss {loss[0]:.3e} -> {final[-1]:.3e}")
    tr.close()
    return P, {"runs": nruns, "box": [nxs, nys, nzs], "nt": nt, "offset": off, "counts": counts.tolist(),
               "final_loss": final, "steps": steps, "batch": batch, "noise": noise, "seconds": time.time() - t0}


def _inner(a, nxs, nys, nzs, off=0):
    return a[off + 3:off + 3 + nys, off + 3:off + 3 + nxs, 3:3 + nzs]


def evaluate_boundary(params, seed=101, nxs=64, nys=64, nzs=48, dx=0.25, dt=0.01, nt=500, every=20, zbeg=-0.75,
                      layers=problem.LHM_EXAMPLE, margin=6, extra=(20,), eps=0.05, sponges=(SPONGE,)):
    """One held-out case run several ways on the GPU, measured against the
    reference run (the truth) over one region U: the inner box minus
    ``margin`` cells at each absorbing face.
      lab      the inner box with the learned boundary (na = 0) and the
               default strip sponge; for each other s in sponges also
               lab_s<s>, the same with sponge strength s
      zero     the same grid with zero ghosts (the boundary without a net)
      pml6     the same grid with a 6-cell ADE-PML (its interior is U)
      pml<n>   for n in extra: the box grown so that an n-cell PML leaves
               the same interior U
    Returns per method: rms_rel = sqrt(sum_t |V - V_true|^2 / sum_t
    |V_true|^2) over U at every ``every`` steps, peak_rel = max_t |V -
    V_true| / max_t |V_true|, final_rel at the last sample, trace_misfit =
    relative L2 of 16 surface receivers' velocity traces, cells and
    ms_per_step."""
    import time
    from .solver import Swpc3D
    lc = wide_case(nxs, nys, nzs, dx, dt, nt, seed, zbeg=zbeg, layers=layers, eps=eps)
    wst, _ = station_cells3(lc.wide.grid, lc.wide.medium.kfs, lc.wide.medium.kob, _stations(lc, 4))
    runs = {"truth": (lc.wide, None, 0, wst)}
    for sp in sponges:
        c = inner_case(lc, 0, nst_side=4)
        runs["lab" if sp == SPONGE else f"lab_s{sp:g}"] = (c, (params, sp), 0, c.stations)
    c = inner_case(lc, 0, nst_side=4)
    runs["zero"] = (c, None, 0, c.stations)
    c = inner_case(lc, margin, nst_side=4)
    runs[f"pml{margin}"] = (c, None, 0, c.stations)
    for na in extra:
        c = inner_case(lc, na, nst_side=4, grow=na - margin)
        runs[f"pml{na}"] = (c, None, na - margin, c.stations)
    sims, grow = {}, {}
    for k, (cs, p, gr, st) in runs.items():
        sims[k] = Swpc3D(cs.grid, cs.medium, cs.pml, cs.glut, st, 1, nt)
        if p is not None:
            sims[k].set_learned_boundary(p[0])
            sims[k].set_learned_boundary_sponge(p[1])
        grow[k] = lc.off if k == "truth" else gr
    def region(k, f):
        o = grow[k]
        a = sims[k].field(f)
        return a[o + 3 + margin:o + 3 + nys - margin, o + 3 + margin:o + 3 + nxs - margin, 3:3 + nzs - margin]
    err = {k: [] for k in runs if k != "truth"}
    pw = []
    it = 1
    while it <= nt:
        n = min(every, nt - it + 1)
        for s in sims.values():
            s.run(it, n)
        it += n
        tv = [region("truth", f) for f in ("Vx", "Vy", "Vz")]
        pw.append(sum(float((v ** 2).sum()) for v in tv))
        for k in err:
            fv = [region(k, f) for f in ("Vx", "Vy", "Vz")]
            err[k].append(sum(float(((a_ - b_) ** 2).sum()) for a_, b_ in zip(fv, tv)))
    pw = np.array(pw)
    vt = sims["truth"].receivers()[0].astype(np.float64)
    out = {}
    for k, e in err.items():
        e = np.array(e)
        v = sims[k].receivers()[0].astype(np.float64)
        g = runs[k][0].grid
        sims[k].sync()
        t0 = time.perf_counter()
        sims[k].run(1, 50)
        sims[k].sync()
        ms = (time.perf_counter() - t0) / 50 * 1e3
        out[k] = {"rms_rel": float(np.sqrt(e.sum() / pw.sum())), "peak_rel": float(np.sqrt(e.max() / pw.max())),
                  "final_rel": float(np.sqrt(e[-1] / max(pw[-1], 1e-300))),
                  "trace_misfit": float(np.linalg.norm(v - vt) / max(np.linalg.norm(vt), 1e-300)),
                  "cells": int(g.nx * g.ny * g.nz), "ms_per_step": ms}
    for s in sims.values():
        s.close()
    return out


def _main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(prog="python -m openswpc_amd.lab",
                                 description="train or evaluate the learned absorbing boundary on the GPU")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--runs", type=int, default=200)
    t.add_argument("--steps", type=int, default=6000)
    t.add_argument("--box", type=int, nargs=3, default=(64, 64, 48), help="inner box of the reference runs (cells)")
    t.add_argument("--nt", type=int, default=500)
    t.add_argument("--offset", type=int, default=OFFSET,
                   help="cells between the inner box and the reflecting edges (more than half the distance the "
                        "fastest wave covers in nt steps)")
    t.add_argument("--noise", type=float, default=0.1)
    t.add_argument("--seed", type=int, default=1)
    t.add_argument("--out", default=str(WEIGHTS))
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=str(WEIGHTS))
    e.add_argument("--seed", type=int, default=101)
    e.add_argument("--nt", type=int, default=500)
    a = ap.parse_args(argv)
    if a.cmd == "train":
        nxs, nys, nzs = a.box
        P, rep = train_boundary(nruns=a.runs, nxs=nxs, nys=nys, nzs=nzs, nt=a.nt, off=a.offset, steps=a.steps,
                                noise=a.noise, seed=a.seed)
        save_weights(a.out, P, report=json.dumps(rep))
        print(json.dumps(rep))
        print(json.dumps(evaluate_boundary(P, seed=10_001)))
    else:
        P, _ = load_weights(a.weights)
        print(json.dumps(evaluate_boundary(P, seed=a.seed, nt=a.nt)))


if __name__ == "__main__":
    _main()


def interior_case(case: problem.Case3D):
    """A case with a PML of na cells restated as the learned-boundary grid
    that covers the same physical region: its interior box (nx - 2na) x
    (ny - 2na) x (nz - na) with na = 0, the same assembled medium, the
    source cells shifted by the absorber width and the receivers at the
    same physical positions."""
    from .source import SourceTable
    g0, m0 = case.grid, case.medium
    na = g0.na
    g = md.Grid3D(g0.nx - 2 * na, g0.ny - 2 * na, g0.nz - na, g0.dx, g0.dy, g0.dz, g0.dt, na=0,
                  xbeg=float(np.float32(g0.xbeg + na * np.float32(g0.dx))),
                  ybeg=float(np.float32(g0.ybeg + na * np.float32(g0.dy))), zbeg=g0.zbeg)
    sj, si, sk = slice(na, na + g.ny + 6), slice(na, na + g.nx + 6), slice(0, g.nz + 6)
    kw = {n: np.ascontiguousarray(getattr(m0, n)[sj, si, sk])
          for n in ("rho", "lam", "mu", "taup", "taus", "bx", "by", "bz", "muyz", "muxz", "muxy")}
    kw.update({n: np.ascontiguousarray(getattr(m0, n)[sj, si])
               for n in ("kfs", "kob", "kfs_top", "kfs_bot", "kob_top", "kob_bot")})
    med = md.Medium3D(**kw, ts=m0.ts, zeta=m0.zeta, vmin=m0.vmin, vmax=m0.vmax)
    gl = case.glut
    glut = SourceTable(np.ascontiguousarray(gl.idx - np.array([na, na, 0], np.int32)), gl.m, gl.table, gl.M0,
                       gl.fcut)
    cells, kept = station_cells3(g, med.kfs, med.kob, case.station_list)
    if len(kept) != len(case.station_list):
        raise ValueError("a receiver falls outside the interior box")
    prof = tuple(np.zeros((n, 4), np.float32) for n in (g.nx, g.nx, g.ny, g.ny, g.nz, g.nz))
    return problem.Case3D(g, med, prof, glut, cells, kept, case.nt, case.ntdec_w, case.mech)
