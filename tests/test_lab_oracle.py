"""Learned absorbing boundary: the oracle pinned on the CPU.

The FP64 network gradient is pinned by central finite differences; the C
forward against the FP64 one; the ghost fill's feature layout, target
cells, stress scaling, previous-ghost recurrence and air-column rule by
nets whose only non-zero weights copy one chosen feature (so the expected
ghost value is known in closed form); the training gather against the
ghost fill's own features and the wide grid's ghost cells.
"""
import ctypes as C

import numpy as np
import pytest

from openswpc_amd import lab, medium as md, problem
from oracle import lab as olab
from oracle.oracle import Oracle3D, olib

F = ["Vx", "Vy", "Vz", "Sxx", "Syy", "Szz", "Syz", "Sxz", "Sxy"]


def _tiny_case(nx=14, ny=12, nz=10, zbeg=-0.5, seed=3):
    g = md.Grid3D(nx, ny, nz, 0.25, 0.25, 0.25, 0.01, na=0, zbeg=zbeg)
    rho, lam, mu, qp, qs = md.layered3(g, problem.LHM_EXAMPLE)
    med = md.medium_setup3(g, rho, lam, mu, qp, qs)
    prof = tuple(np.zeros((n, 4), np.float32) for n in (nx, nx, ny, ny, nz, nz))
    return problem.Case3D(g, med, prof, None, np.zeros((0, 3), np.int32), [], 10, 1, (0, 0, 0))


def _random_fields(o, seed):
    rng = np.random.default_rng(seed)
    for n in F:
        o.f[n][...] = rng.standard_normal(o.f[n].shape).astype(np.float32)


def test_layout_matches_product():
    assert olab.NP == lab.NP == olib(32).oracle_lab_nparam()
    assert olab.offsets() == lab.offsets()
    for net in range(10):
        assert olib(32).oracle_lab_nout(net) == lab.NOUT[net]


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.standard_normal(olab.NP) * 0.3
    X = rng.standard_normal((40, olab.NIN))
    T = rng.standard_normal((40, olab.NO))
    for nout in (4, 5):
        loss, g = olab.loss_grad(p, X, T, nout)
        assert abs(loss - olab.loss_only(p, X, T, nout)) < 1e-12 * max(1.0, loss)
        o = olab.offsets()
        # parameters from every block, including the unused 5th output row when nout = 4
        idx = np.concatenate([rng.integers(o[a], o[b], 8) for a, b in
                              (("W1", "b1"), ("b1", "W2"), ("W2", "b2"), ("b2", "W3"), ("W3", "b3"), ("b3", "Ws"),
                               ("Ws", "n"))])
        h = 1e-6
        for q in idx:
            pp, pm = p.copy(), p.copy()
            pp[q] += h
            pm[q] -= h
            fd = (olab.loss_only(pp, X, T, nout) - olab.loss_only(pm, X, T, nout)) / (2 * h)
            assert abs(fd - g[q]) <= 1e-7 + 1e-6 * abs(fd), (nout, q, fd, g[q])


def test_c_forward_matches_fp64():
    """The C forward (BF16 inputs to the hidden layers) against the FP64
    one with the same roundings; and the rounding is what separates it from
    the plain FP64 net (by about 2^-9 of the hidden path)."""
    rng = np.random.default_rng(1)
    p = olab.bf16(rng.standard_normal(olab.NP) * 0.2).astype(np.float32)
    L = olib(32)
    far = 0.0
    for _ in range(20):
        x = rng.standard_normal(olab.NIN).astype(np.float32)
        y = np.zeros(5, np.float32)
        L.oracle_lab_forward(p.ctypes.data_as(C.POINTER(C.c_float)), x.ctypes.data_as(C.POINTER(C.c_float)),
                             y.ctypes.data_as(C.POINTER(C.c_float)))
        ref = olab.forward(p, x[None], hidden_bf16=True)[0]
        assert np.allclose(y, ref, rtol=1e-5, atol=1e-5)
        far = max(far, float(np.abs(olab.forward(p, x[None])[0] - ref).max()))
    assert 1e-4 < far < 0.1


def _target(net, q):
    f, h = C.c_int(), C.c_int()
    olib(32).oracle_lab_target(net, q, C.byref(f), C.byref(h))
    return f.value, h.value


def _strip_index(face, d, n):
    """padded index along the face axis of strip cell d (1 outermost) or
    ghost depth h at d = 1 - h"""
    return (d if face in (0, 2) else n + 1 - d) + 2


def _sl(o, face, d):
    g = o.case.grid
    n = (g.nx, g.nx, g.ny, g.ny, g.nz)[face]
    c = _strip_index(face, d, n)
    J, I, K = slice(3, g.ny + 3), slice(3, g.nx + 3), slice(3, g.nz + 3)
    if face < 2:
        return (J, c, K)
    if face < 4:
        return (c, I, K)
    return (J, I, c)


@pytest.mark.parametrize("phase", [0, 1])
def test_fill_copy_nets_pin_the_layout(phase):
    """A net whose only weight is Ws[q][field*6 + d-1] = 1 writes, into its
    ghost q, the value of that field at strip cell d: exact up to the
    rounding of x/N*N (and of /imp*imp for stresses)."""
    case = _tiny_case()
    o = Oracle3D(case, 32)
    _random_fields(o, 10 + phase)
    P = np.zeros((10, olab.NP), np.float32)
    picks = {}
    for net in range(phase, 10, 2):
        for q in range(lab.NOUT[net]):
            fi, h = _target(net, q)
            d = 1 + (q % 3)
            fsrc = (fi + q) % 9            # a different field than the target for some q
            _, _, _, _, _, _, Ws = olab.unpack(P[net])
            Ws[q, fsrc * 6 + d - 1] = 1.0
            picks[(net, q)] = (fi, h, fsrc, d)
    o.set_lab(P, sponge=0.0)
    before = {n: o.f[n].copy() for n in F}
    o.lab_fill(phase)
    g = case.grid
    air = ~((case.medium.lam + 2 * case.medium.mu) > 0)
    rho = case.medium.rho
    vp = np.sqrt(np.where(air, 1.0, case.medium.lam + 2 * case.medium.mu) / rho)
    imp = np.where(air, 1.0, rho * vp)
    for (net, q), (fi, h, fsrc, d) in picks.items():
        face = net // 2
        got = o.f[F[fi]][_sl(o, face, 1 - h)]
        want = before[F[fsrc]][_sl(o, face, d)].astype(np.float64)
        im1 = imp[_sl(o, face, 1)]
        want = want * np.where(fsrc >= 3, 1.0 / im1, 1.0) * np.where(fi >= 3, im1, 1.0)
        solid = ~air[_sl(o, face, 1)]
        want = np.where(solid, want, 0.0)
        assert np.allclose(got, want, rtol=3e-6, atol=1e-6), (net, q)
        assert (got[~solid] == 0).all()


def test_fill_previous_ghost_recurrence():
    """Ws[q][54 + q] = 1 makes every ghost keep its previous value."""
    case = _tiny_case()
    o = Oracle3D(case, 32)
    _random_fields(o, 7)
    P = np.zeros((10, olab.NP), np.float32)
    for net in range(10):
        _, _, _, _, _, _, Ws = olab.unpack(P[net])
        for q in range(lab.NOUT[net]):
            Ws[q, 54 + q] = 1.0
    o.set_lab(P, sponge=0.0)
    air = ~((case.medium.lam + 2 * case.medium.mu) > 0)
    for phase in (0, 1):
        before = {n: o.f[n].copy() for n in F}
        o.lab_fill(phase)
        for net in range(phase, 10, 2):
            face = net // 2
            solid = ~air[_sl(o, face, 1)]
            for q in range(lab.NOUT[net]):
                fi, h = _target(net, q)
                got = o.f[F[fi]][_sl(o, face, 1 - h)]
                want = np.where(solid, before[F[fi]][_sl(o, face, 1 - h)], 0.0)
                assert np.allclose(got, want, rtol=3e-6, atol=1e-6)


@pytest.mark.parametrize("phase", [0, 1])
def test_fill_sponge_profile(phase):
    """With all-zero nets the fill writes zero ghosts and the sponge alone
    acts: every interior cell of the phase's own fields (V at 0, S at 1)
    is multiplied, once per absorbing face whose strip holds it (depth d =
    1..6 from that face), by the FP32 factor 1 - s ((7 - d)/6)^2; the other
    phase's fields and the cells deeper than 6 are untouched."""
    s = 0.3
    case = _tiny_case(nx=16, ny=15, nz=14)
    o = Oracle3D(case, 64)
    _random_fields(o, 20 + phase)
    o.set_lab(np.zeros((10, olab.NP), np.float32), sponge=s)
    before = {n: o.f[n].copy() for n in F}
    o.lab_fill(phase)
    g = case.grid
    i = np.arange(1, g.nx + 1)[None, :, None]
    j = np.arange(1, g.ny + 1)[:, None, None]
    k = np.arange(1, g.nz + 1)[None, None, :]
    fac = np.ones((g.ny, g.nx, g.nz))
    for depth in (i, g.nx + 1 - i, j, g.ny + 1 - j, g.nz + 1 - k):
        r = np.float32(1.0 - s * ((7 - depth) / 6.0) ** 2).astype(np.float64)
        fac = fac * np.where(depth <= 6, r, 1.0)
    assert (fac < 1).sum() > 0 and (fac == 1).sum() > 0
    inner = (slice(3, g.ny + 3), slice(3, g.nx + 3), slice(3, g.nz + 3))
    for n in range(9):
        got = o.f[F[n]][inner].astype(np.float64)
        was = before[F[n]][inner].astype(np.float64)
        own = (n < 3) == (phase == 0)
        if own:
            assert np.allclose(got, was * fac, rtol=1e-6, atol=0), F[n]
        else:
            assert (got == was).all(), F[n]


def test_gather_records_match_fill_features_and_ghosts():
    """Gathering the inner box of a grid: the targets are that box's ghost
    values (/N, stresses /rho*vp), and a copy net applied through the fill
    of a grid equal to the box reproduces target = copied feature * N."""
    case = _tiny_case(nx=20, ny=18, nz=14)
    o = Oracle3D(case, 32)
    _random_fields(o, 5)
    oi, oj, nxs, nys, nzs = 3, 2, 12, 11, 10
    o.lab_gather_begin(oi, oj, nxs, nys, nzs)
    rec = o.lab_gather(0)
    g = case.grid
    rho, lam, mu = case.medium.rho, case.medium.lam, case.medium.mu
    for face in range(5):
        r = rec[face]
        assert r.shape[1] == olab.REC and len(r) > 0
        ext = (nys, nys, nxs, nxs, nxs)[face]
        n = (nxs, nxs, nys, nys, nzs)[face]
        for row in r[:: max(1, len(r) // 25)]:
            col = int(row[66])
            a, b = col % ext + 1, col // ext + 1
            N, im = row[67], row[68]
            for q in range(lab.NOUT[2 * face]):
                fi, h = _target(2 * face, q)
                p = (1 - h) if face in (0, 2) else n + h
                if face < 2:
                    i, j, k = p, a, b
                elif face < 4:
                    i, j, k = a, p, b
                else:
                    i, j, k = a, b, p
                v = float(o.f[F[fi]][j + oj + 2, i + oi + 2, k + 2])
                if fi >= 3:
                    v /= im
                assert abs(row[61 + q] - v / N) <= 1e-5 * (1 + abs(v / N))
            # the strip features of the outermost cell, field Vx
            p1 = 1 if face in (0, 2) else n
            i, j, k = (p1, a, b) if face < 2 else ((a, p1, b) if face < 4 else (a, b, p1))
            c = (j + oj + 2, i + oi + 2, k + 2)
            assert abs(row[0] - o.f["Vx"][c] / N) <= 1e-5 * (1 + abs(row[0]))
            l2m = lam[c] + 2 * mu[c]
            vp = np.sqrt(l2m / rho[c])
            assert abs(im - rho[c] * vp) <= 1e-5 * im
            assert abs(row[59] - np.sqrt(mu[c] / rho[c]) / vp) < 1e-5
            h_n = (g.dx, g.dx, g.dy, g.dy, g.dz)[face]
            assert abs(row[60] - vp * g.dt / h_n) < 1e-5
    # a second call carries this call's targets as previous-ghost features
    rec2 = o.lab_gather(0)
    for face in range(5):
        r1, r2 = rec[face], rec2[face]
        m1 = {int(x[66]): x for x in r1}
        for row in r2[:: max(1, len(r2) // 25)]:
            x1 = m1[int(row[66])]
            prev = x1[61:66] * x1[67]           # raw previous ghosts
            assert np.allclose(row[54:59] * row[67], prev, rtol=1e-5, atol=1e-6)
