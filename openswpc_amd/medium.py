"""Grid, velocity models and medium assembly (the setup-time producers of
every coefficient array the kernels read).

Restates, on the host in numpy float32:
  * the grid and region bounds of ``global__setup2``
    (src/swpc_3d/m_global.f90:221-390; src/swpc_psv/m_global.f90:201-310)
  * layered and layered-plus-random velocity models
    (src/swpc_3d/m_vmodel_lhm.f90:21-169, m_vmodel_lhm_rmed.f90:23-229,
    m_vmodel_uni.f90) and the periodic tiling of a random-medium volume
    (src/shared/m_rdrmed.f90:73-134)
  * ``medium__setup`` (src/swpc_3d/m_medium.f90:36-467): absorber
    homogenisation, tau-method Q, unrelaxed moduli at fq_ref, free-surface /
    ocean-bottom detection and the staggered-grid averages; and its P-SV
    twin (src/swpc_psv/m_medium.f90).

Arrays follow the reference's Fortran layout: a 3-D field ``a(k,i,j)``
over the memory box [-2, n+3]^3 is a C-ordered numpy array of shape
(ny+6, nx+6, nz+6) indexed ``a[j+2, i+2, k+2]``; a P-SV field ``a(k,i)``
has shape (nx+6, nz+6).
"""
from __future__ import annotations

from dataclasses import dataclass, field
import math

import numpy as np

from . import fdtool

F32 = np.float32
EPS = np.finfo(np.float32).eps
NM = 3  # m_global.f90:31


# ---------------------------------------------------------------------------
# grid
# ---------------------------------------------------------------------------
@dataclass
class Grid3D:
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    dt: float
    na: int = 20
    xbeg: float | None = None
    ybeg: float | None = None
    zbeg: float | None = None
    tbeg: float = 0.0

    def __post_init__(self):
        # defaults of global__readprm (m_global.f90:167-169)
        if self.xbeg is None:
            self.xbeg = float(F32(-(self.nx // 2) * F32(self.dx)))
        if self.ybeg is None:
            self.ybeg = float(F32(-(self.ny // 2) * F32(self.dy)))
        if self.zbeg is None:
            self.zbeg = float(F32(-30 * F32(self.dz)))
        self.xend = float(F32(F32(self.xbeg) + self.nx * F32(self.dx)))
        self.yend = float(F32(F32(self.ybeg) + self.ny * F32(self.dy)))
        self.zend = float(F32(F32(self.zbeg) + self.nz * F32(self.dz)))
        self.xc = fdtool.centres(-2, self.nx + 3, self.xbeg, self.dx)
        self.yc = fdtool.centres(-2, self.ny + 3, self.ybeg, self.dy)
        self.zc = fdtool.centres(-2, self.nz + 3, self.zbeg, self.dz)

    @property
    def shape(self):
        return (self.ny + 6, self.nx + 6, self.nz + 6)

    def zeros(self, dtype=np.float32):
        return np.zeros(self.shape, dtype=dtype)

    def kbeg_a(self):
        """PML column start (m_global.f90:336-345) over i, j in [1, n]."""
        i = np.arange(1, self.nx + 1)[None, :]
        j = np.arange(1, self.ny + 1)[:, None]
        side = (i <= self.na) | (i >= self.nx - self.na + 1) | (j <= self.na) | (j >= self.ny - self.na + 1)
        return np.where(side, 1, self.nz - self.na + 1).astype(np.int32)

    def n_pml_cells(self):
        return int((self.nz - self.kbeg_a() + 1).sum())

    def n_interior_cells(self):
        return (self.nx - 2 * self.na) * (self.ny - 2 * self.na) * (self.nz - self.na)


@dataclass
class Grid2D:
    nx: int
    nz: int
    dx: float
    dz: float
    dt: float
    na: int = 20
    xbeg: float | None = None
    zbeg: float | None = None
    tbeg: float = 0.0

    def __post_init__(self):
        if self.xbeg is None:
            self.xbeg = float(F32(-(self.nx // 2) * F32(self.dx)))
        if self.zbeg is None:
            self.zbeg = float(F32(-30 * F32(self.dz)))
        self.xend = float(F32(F32(self.xbeg) + self.nx * F32(self.dx)))
        self.zend = float(F32(F32(self.zbeg) + self.nz * F32(self.dz)))
        self.xc = fdtool.centres(-2, self.nx + 3, self.xbeg, self.dx)
        self.zc = fdtool.centres(-2, self.nz + 3, self.zbeg, self.dz)

    @property
    def shape(self):
        return (self.nx + 6, self.nz + 6)

    def zeros(self, dtype=np.float32):
        return np.zeros(self.shape, dtype=dtype)

    def n_pml_cells(self):
        return 2 * self.na * self.nz + (self.nx - 2 * self.na) * self.na

    def n_interior_cells(self):
        return (self.nx - 2 * self.na) * (self.nz - self.na)


# ---------------------------------------------------------------------------
# velocity models
# ---------------------------------------------------------------------------
@dataclass
class Layer:
    depth: float
    rho: float
    vp: float
    vs: float
    qp: float
    qs: float


def read_lhm(path):
    """Layer file of vmodel_lhm: depth rho vp vs qp qs [rmed file] per line."""
    out = []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            w = s.split()
            out.append(Layer(*[float(x) for x in w[:6]]))
    return out


def _vcut(layers, vcut):
    # velocity cut-off loop of vmodel_lhm (m_vmodel_lhm.f90:104-113)
    L = [Layer(**vars(l)) for l in layers]
    for l in range(len(L) - 2, -1, -1):
        if (L[l].vp < vcut or L[l].vs < vcut) and (L[l].vp > 0 and L[l].vs > 0):
            L[l] = Layer(L[l].depth, L[l + 1].rho, L[l + 1].vp, L[l + 1].vs, L[l + 1].qp, L[l + 1].qs)
    return L


def _seawater_vel(z, munk=False):
    # seawater__vel (src/shared/m_seawater.f90:29-47): 1.5 km/s, or the Munk
    # profile with epsilon 0.00737 when munk_profile is set
    eps = 0.00737 if munk else 0.0
    zb = 2 * (float(z) * 1000.0 - 1300.0) / 1300.0
    return F32(1.5 * (1.0 + eps * (zb - 1.0 + math.exp(-zb))))


def _layer_index(zc, depths):
    """Last layer l with zc >= depth(l) (m_vmodel_lhm.f90:132-141), -1 above."""
    idx = np.full(zc.shape, -1, dtype=np.int64)
    for l, d in enumerate(depths):
        idx = np.where(zc >= F32(d), l, idx)
    return idx


def vmodel_lhm(zc, layers, vcut=0.0, xi=None, vmax_clip=None, rhomin=1.0, munk=False):
    """1-D layered model over the k axis of the memory box.

    Returns per-k arrays (rho, lam, mu, qp, qs) when ``xi`` is None; with a
    random field ``xi`` (shape (..., nk)) returns full arrays with the
    perturbation of vmodel_lhm_rmed: rho(1+0.8 xi), vp(1+xi), vs(1+xi) and
    the ``vcheck`` clamp against (vcut, vmax_clip, rhomin).
    """
    L = _vcut(layers, vcut)
    depths = [l.depth for l in L]
    li = _layer_index(zc, depths)
    rho0 = np.array([l.rho for l in L], F32)
    vp0 = np.array([l.vp for l in L], F32)
    vs0 = np.array([l.vs for l in L], F32)
    qp0 = np.array([l.qp for l in L], F32)
    qs0 = np.array([l.qs for l in L], F32)
    above = zc < F32(depths[0])
    air = above & (zc < 0)
    ocean = above & ~air
    lc = np.clip(li, 0, None)
    if xi is None:
        rho = rho0[lc].copy()
        vp = vp0[lc].copy()
        vs = vs0[lc].copy()
        mu = (rho * vs * vs).astype(F32)
        lam = (rho * (vp * vp - 2 * vs * vs)).astype(F32)
        qp = qp0[lc].copy()
        qs = qs0[lc].copy()
        shape_mask = lambda m: m
    else:
        xi = np.asarray(xi, F32)
        rho = (rho0[lc] * (1 + F32(0.8) * xi)).astype(F32)
        vp = (vp0[lc] * (1 + xi)).astype(F32)
        vs = (vs0[lc] * (1 + xi)).astype(F32)
        solid = (vp0[lc] > 0) & (vs0[lc] > 0) & ~above
        vmin = F32(vcut)
        vmax = F32(vmax_clip if vmax_clip is not None else 1e30)
        vp2, vs2, rho2, _ = fdtool.vcheck(vp, vs, rho, xi, vmin, vmax, rhomin)
        vp = np.where(solid, vp2, vp).astype(F32)
        vs = np.where(solid, vs2, vs).astype(F32)
        rho = np.where(solid, rho2, rho).astype(F32)
        mu = (rho * vs * vs).astype(F32)
        lam = (rho * (vp * vp - 2 * vs * vs)).astype(F32)
        qp = np.broadcast_to(qp0[lc], xi.shape).astype(F32)
        qs = np.broadcast_to(qs0[lc], xi.shape).astype(F32)
        air = np.broadcast_to(air, xi.shape)
        ocean = np.broadcast_to(ocean, xi.shape)
    rho = np.where(air, F32(0.001), rho).astype(F32)
    mu = np.where(air, F32(0.0), mu).astype(F32)
    lam = np.where(air, F32(0.0), lam).astype(F32)
    qp = np.where(air, F32(10.0), qp).astype(F32)
    qs = np.where(air, F32(10.0), qs).astype(F32)
    if np.any(ocean):
        zz = np.broadcast_to(zc, ocean.shape)
        vw = np.vectorize(lambda z: _seawater_vel(z, munk), otypes=[F32])(zz)
        rho = np.where(ocean, F32(1.0), rho).astype(F32)
        mu = np.where(ocean, F32(0.0), mu).astype(F32)
        lam = np.where(ocean, vw * vw, lam).astype(F32)
        qp = np.where(ocean, F32(1e6), qp).astype(F32)
        qs = np.where(ocean, F32(1e6), qs).astype(F32)
    return rho, lam, mu, qp, qs


def vmodel_uni(zc, shape, vp0=5.0, vs0=None, rho0=2.7, qp0=1e6, qs0=1e6, topo0=0.0, xi=None, vmin=0.0,
               vmax_clip=None, rhomin=1.0, munk=False):
    """vmodel_uni / vmodel_uni_rmed (src/swpc_3d/m_vmodel_uni.f90:22-146,
    m_vmodel_uni_rmed.f90:22-175): a half-space below z = topo0 (strictly
    deeper cells), ocean between 0 and topo0, air above; with a random field
    xi the solid takes rho(1+0.8 xi), v(1+xi) and the vcheck clamp.  vs0
    defaults to vp0/sqrt(3) in real(SP).  Returns full arrays of ``shape``
    (the k axis last)."""
    vp0 = F32(vp0)
    vs0 = F32(vp0 / F32(math.sqrt(3.0))) if vs0 is None else F32(vs0)
    rho0, qp0, qs0, topo0 = F32(rho0), F32(qp0), F32(qs0), F32(topo0)
    z = np.broadcast_to(np.asarray(zc, F32), shape)
    solid = z > topo0
    ocean = ~solid & (z > 0)
    if xi is None:
        rho_s = np.full(shape, rho0, F32)
        vp_s = np.full(shape, vp0, F32)
        vs_s = np.full(shape, vs0, F32)
    else:
        xi = np.asarray(xi, F32)
        rho_s = ((F32(1.0) + F32(0.8) * xi) * rho0).astype(F32)
        vp_s = ((F32(1.0) + xi) * vp0).astype(F32)
        vs_s = ((F32(1.0) + xi) * vs0).astype(F32)
        vp_s, vs_s, rho_s, _ = fdtool.vcheck(vp_s, vs_s, rho_s, xi, F32(vmin),
                                             F32(vmax_clip if vmax_clip is not None else 1e30), rhomin)
    rho = np.where(solid, rho_s, np.where(ocean, F32(1.0), F32(0.001))).astype(F32)
    mu = np.where(solid, rho_s * vs_s * vs_s, F32(0.0)).astype(F32)
    vw = np.vectorize(lambda zz: _seawater_vel(zz, munk), otypes=[F32])(np.asarray(zc, F32))
    vw = np.broadcast_to(vw, shape)
    lam = np.where(solid, rho_s * (vp_s * vp_s - F32(2) * vs_s * vs_s),
                   np.where(ocean, F32(1.0) * (vw * vw), F32(0.0))).astype(F32)
    qp = np.where(solid, qp0, np.where(ocean, F32(1e6), F32(10.0))).astype(F32)
    qs = np.where(solid, qs0, np.where(ocean, F32(1e6), F32(10.0))).astype(F32)
    return rho, lam, mu, qp, qs


def read_lhm_rmed(path):
    """Layer file of vmodel_lhm_rmed (m_vmodel_lhm_rmed.f90:101-110):
    depth rho vp vs qp qs rmed-file per line -> (layers, file names)."""
    layers, files = [], []
    with open(path) as f:
        for line in f:
            t = line.strip()
            if not t or t.startswith("#"):
                continue
            w = t.split()
            layers.append(Layer(*[float(x) for x in w[:6]]))
            files.append(w[6] if len(w) > 6 else "")
    return layers, files


def layer_xi(zc, layers, vcut, fields):
    """The random field each cell of vmodel_lhm_rmed perturbs with: the
    field of the deepest layer whose top lies at or above the cell
    (m_vmodel_lhm_rmed.f90:176-196).  ``fields[l]`` is a full array or None
    (no file: zero)."""
    L = _vcut(layers, vcut)
    li = _layer_index(np.asarray(zc, F32), [l.depth for l in L])
    shape = next(f.shape for f in fields if f is not None)
    out = np.zeros(shape, F32)
    for l, f in enumerate(fields):
        if f is not None:
            sel = np.broadcast_to(li == l, shape)
            out = np.where(sel, f, out)
    return out


def rmed_tile2(vol, i0, i1, k0, k1):
    """Periodic tiling of a 2-D random medium vol[nxc, nzc] onto
    [i0,i1] x [k0,k1] (src/shared/m_rdrmed.f90:24-70); returns (ni, nk)."""
    nxc, nzc = vol.shape
    ii = np.arange(i0, i1 + 1) % nxc
    ii = np.where(ii <= 0, ii + nxc, ii) - 1
    kk = np.arange(k0, k1 + 1) % nzc
    kk = np.where(kk <= 0, kk + nzc, kk) - 1
    return np.ascontiguousarray(vol[ii][:, kk])


def rmed_tile(vol, i0, i1, j0, j1, k0, k1):
    """Periodic tiling of a random-medium volume vol[nxc, nyc, nzc] onto the
    index box [i0,i1]x[j0,j1]x[k0,k1] (src/shared/m_rdrmed.f90:111-132).
    Returns an array of shape (nj, ni, nk) in (j, i, k) order."""
    nxc, nyc, nzc = vol.shape
    ii = (np.arange(i0, i1 + 1) % nxc)
    ii = np.where(ii <= 0, ii + nxc, ii) - 1
    jj = (np.arange(j0, j1 + 1) % nyc)
    jj = np.where(jj <= 0, jj + nyc, jj) - 1
    kk = np.arange(k0, k1 + 1)
    kk = np.where(kk <= 0, kk + nzc, kk)
    kk = kk % nzc
    kk = np.where(kk <= 0, kk + nzc, kk) - 1
    return np.ascontiguousarray(vol[ii][:, jj][:, :, kk].transpose(1, 0, 2))


# ---------------------------------------------------------------------------
# medium__setup
# ---------------------------------------------------------------------------
@dataclass
class Medium3D:
    rho: np.ndarray
    lam: np.ndarray
    mu: np.ndarray
    taup: np.ndarray
    taus: np.ndarray
    bx: np.ndarray
    by: np.ndarray
    bz: np.ndarray
    muyz: np.ndarray
    muxz: np.ndarray
    muxy: np.ndarray
    kfs: np.ndarray
    kob: np.ndarray
    kfs_top: np.ndarray
    kfs_bot: np.ndarray
    kob_top: np.ndarray
    kob_bot: np.ndarray
    ts: np.ndarray
    zeta: float
    vmin: float
    vmax: float


def _homogenize3(a, g: Grid3D):
    """Absorber homogenisation of m_medium.f90:124-190 (in place)."""
    na, nx, ny, nz = g.na, g.nx, g.ny, g.nz
    o = 2
    a[:, 0:na + o + 1, :] = a[:, na + 1 + o:na + 2 + o, :]
    a[:, nx - na + 1 + o:, :] = a[:, nx - na + o:nx - na + o + 1, :]
    a[0:na + o + 1, :, :] = a[na + 1 + o:na + 2 + o, :, :]
    a[ny - na + 1 + o:, :, :] = a[ny - na + o:ny - na + o + 1, :, :]
    a[:, :, nz - na + 1 + o:] = a[:, :, nz - na + o:nz - na + o + 1]


def _relax3(rho, lam, mu, taup, taus, ts, fq_ref):
    """relaxed_medium (m_medium.f90:234-274): unrelaxed moduli at fq_ref."""
    cc = np.complex64(fdtool.relaxed_cc(len(ts), ts, fq_ref))
    one = np.complex64(1.0)
    rho_beta2 = mu.copy()
    rho_alpha2 = (lam + 2 * mu).astype(F32)
    chi_mu = (F32(1.0) / np.real(one / np.sqrt(one - cc * taus.astype(np.complex64)))).astype(F32)
    chi_lam = (F32(1.0) / np.real(one / np.sqrt(one - cc * taup.astype(np.complex64)))).astype(F32)
    mu_n = (rho_beta2 / chi_mu ** 2).astype(F32)
    lam_n = (rho_alpha2 / chi_lam ** 2 - 2 * mu_n).astype(F32)
    return lam_n, mu_n


def _detect(lam, mu, kaxis_last, kbeg, kend, lo_i):
    """kfs/kob: last k in [kbeg, kend-1] with |x(k)| < eps < |x(k+1)|
    (m_medium.f90:357-373).  Arrays are (..., k) with k index offset 2."""
    ks = np.arange(kbeg, kend)  # k = kbeg..kend-1
    def last(x):
        a = np.abs(x[..., ks + 2]) < EPS
        b = np.abs(x[..., ks + 3]) > EPS
        hit = a & b
        anyhit = hit.any(axis=-1)
        lastk = ks[::-1][np.argmax(hit[..., ::-1], axis=-1)]
        return np.where(anyhit, lastk, kbeg - 1).astype(np.int32)
    return last(lam), last(mu)


def medium_setup3(g: Grid3D, rho, lam, mu, qp, qs, fq_min=0.05, fq_max=5.0, fq_ref=1.0, benchmark=False):
    """medium__setup of the reference on memory-box arrays (copies inputs)."""
    rho, lam, mu = rho.astype(F32).copy(), lam.astype(F32).copy(), mu.astype(F32).copy()
    taup, taus = qp.astype(F32).copy(), qs.astype(F32).copy()
    for a in (rho, lam, mu, taup, taus):
        _homogenize3(a, g)
    ts = fdtool.visco_set_relaxtime(NM, fq_min, fq_max)
    zeta = fdtool.visco_constq_zeta(NM, fq_min, fq_max, ts)
    if benchmark:
        zeta = F32(0.0)
    taup = (F32(NM) * zeta / taup).astype(F32)
    taus = (F32(NM) * zeta / taus).astype(F32)
    lam, mu = _relax3(rho, lam, mu, taup, taus, ts, fq_ref)

    # surface detection over i in [0, nx+2], j in [0, ny+2]
    nx, ny, nz = g.nx, g.ny, g.nz
    kfs = np.zeros((ny + 6, nx + 6), np.int32)
    kob = np.zeros((ny + 6, nx + 6), np.int32)
    sl = (slice(0 + 2, ny + 2 + 3), slice(0 + 2, nx + 2 + 3))
    f, o = _detect(lam[sl], mu[sl], True, 1, nz, 0)
    kfs[sl] = f
    kob[sl] = o
    kfs_top = np.zeros_like(kfs)
    kfs_bot = np.zeros_like(kfs)
    kob_top = np.zeros_like(kfs)
    kob_bot = np.zeros_like(kfs)
    from numpy.lib.stride_tricks import sliding_window_view as swv
    # window (i-2:i+3, j-2:j+3) for i, j in [1, n] -> padded indices i..i+5
    # window (i-2:i+3, j-2:j+3) starts at padded index (j, i)
    wf = swv(kfs, (6, 6))[1:ny + 1, 1:nx + 1]
    wo = swv(kob, (6, 6))[1:ny + 1, 1:nx + 1]
    inner = (slice(3, ny + 3), slice(3, nx + 3))
    kfs_top[inner] = np.maximum(wf.min(axis=(2, 3)) - 2, 1)
    kfs_bot[inner] = np.minimum(wf.max(axis=(2, 3)) + 2, nz)
    kob_top[inner] = np.maximum(wo.min(axis=(2, 3)) - 2, 1)
    kob_bot[inner] = np.minimum(wo.max(axis=(2, 3)) + 2, nz)

    # velocity_minmax (m_medium.f90:399-430)
    vmin, vmax = _minmax(lam[2 + 1:ny + 3, 2 + 1:nx + 3, 3:nz + 3], mu[2 + 1:ny + 3, 2 + 1:nx + 3, 3:nz + 3],
                         rho[2 + 1:ny + 3, 2 + 1:nx + 3, 3:nz + 3], kfs[3:ny + 3, 3:nx + 3])

    # averaged_medium (m_medium.f90:432-465) over [1,nx]x[1,ny]x[1,nz]
    bx, by, bz = g.zeros(), g.zeros(), g.zeros()
    muyz, muxz, muxy = g.zeros(), g.zeros(), g.zeros()
    J, I, K = slice(3, ny + 3), slice(3, nx + 3), slice(3, nz + 3)
    J1, I1, K1 = slice(4, ny + 4), slice(4, nx + 4), slice(4, nz + 4)
    two = F32(2.0)
    bx[J, I, K] = two / (rho[J, I, K] + rho[J, I1, K])
    by[J, I, K] = two / (rho[J, I, K] + rho[J1, I, K])
    bz[J, I, K] = two / (rho[J, I, K] + rho[J, I, K1])
    nnn, pnn, npn, ppn = mu[J, I, K], mu[J, I, K1], mu[J, I1, K], mu[J, I1, K1]
    nnp, npp, pnp = mu[J1, I, K], mu[J1, I1, K], mu[J1, I, K1]
    e = F32(EPS)
    muxz[J, I, K] = 4 * nnn * pnn * npn * ppn / (nnn * pnn * npn + nnn * pnn * ppn + nnn * npn * ppn + pnn * npn * ppn + e)
    muxy[J, I, K] = 4 * nnn * npn * nnp * npp / (nnn * npn * nnp + nnn * npn * npp + nnn * nnp * npp + npn * nnp * npp + e)
    muyz[J, I, K] = 4 * nnn * pnn * nnp * pnp / (nnn * pnn * nnp + nnn * pnn * pnp + nnn * nnp * pnp + pnn * nnp * pnp + e)
    return Medium3D(rho, lam, mu, taup, taus, bx, by, bz, muyz, muxz, muxy,
                    kfs, kob, kfs_top, kfs_bot, kob_top, kob_bot, ts, float(zeta), vmin, vmax)


@dataclass
class RawMedium3D:
    """The velocity model's memory-box arrays before medium__setup, for the
    device setup (``Swpc3D`` calls swpc3d_setup_medium with them): the
    absorber homogenisation, Q, relaxation, surface detection and averaged
    coefficients are then formed on the GPU.  ``kfs``, ``kob``, ``vmin``
    and ``vmax`` are filled in by the context that ran the setup."""
    rho: np.ndarray
    lam: np.ndarray
    mu: np.ndarray
    qp: np.ndarray
    qs: np.ndarray
    fq_min: float = 0.05
    fq_max: float = 5.0
    fq_ref: float = 1.0
    benchmark: bool = False
    kfs: np.ndarray = None
    kob: np.ndarray = None
    vmin: float = None
    vmax: float = None

    def constants(self):
        """(ts, zeta3 = NM*zeta as FP32, cc) as medium_setup3 forms them."""
        ts = fdtool.visco_set_relaxtime(NM, self.fq_min, self.fq_max)
        zeta = F32(0.0) if self.benchmark else fdtool.visco_constq_zeta(NM, self.fq_min, self.fq_max, ts)
        zeta3 = F32(F32(NM) * F32(zeta))
        cc = fdtool.relaxed_cc(NM, ts, self.fq_ref)
        return ts, zeta3, cc


def _minmax(lam, mu, rho, kfs):
    """velocity_minmax over k >= kfs+1 of each column."""
    nk = lam.shape[-1]
    k = np.arange(1, nk + 1)
    below = k[None, None, :] >= (kfs[..., None] + 1)
    with np.errstate(divide="ignore", invalid="ignore"):
        vp = np.sqrt((lam + 2 * mu) / rho)
        vs = np.sqrt(mu / rho)
    vmax = float(np.max(np.where(below, vp, -1)))
    vsm = np.where(below & (vs >= EPS), vs, np.inf)
    vmin = float(np.min(vsm)) if np.isfinite(vsm).any() else 1e30
    return vmin, vmax


@dataclass
class Medium2D:
    rho: np.ndarray
    lam: np.ndarray
    mu: np.ndarray
    taup: np.ndarray
    taus: np.ndarray
    kfs: np.ndarray
    kob: np.ndarray
    kfs_top: np.ndarray
    kfs_bot: np.ndarray
    kob_top: np.ndarray
    kob_bot: np.ndarray
    ts: np.ndarray
    zeta: float
    vmin: float
    vmax: float


def medium_setup2(g: Grid2D, rho, lam, mu, qp, qs, fq_min=0.05, fq_max=5.0, fq_ref=1.0, benchmark=False):
    """P-SV medium__setup (src/swpc_psv/m_medium.f90:36-366).

    The reference assigns ``kfs_top`` twice and never ``kfs_bot``
    (src/swpc_psv/m_medium.f90:241-242); this follows the 3-D definition,
    kfs_top = min-2 and kfs_bot = max+2 over the 6-column window, which
    selects the same 2nd-order cells because stresses vanish in the air.
    """
    na, nx, nz = g.na, g.nx, g.nz
    rho, lam, mu = rho.astype(F32).copy(), lam.astype(F32).copy(), mu.astype(F32).copy()
    taup, taus = qp.astype(F32).copy(), qs.astype(F32).copy()
    o = 2
    for a in (rho, lam, mu, taup, taus):
        a[0:na + o + 1, :] = a[na + 1 + o:na + 2 + o, :]
        a[nx - na + 1 + o:, :] = a[nx - na + o:nx - na + o + 1, :]
        a[:, nz - na + 1 + o:] = a[:, nz - na + o:nz - na + o + 1]
    ts = fdtool.visco_set_relaxtime(NM, fq_min, fq_max)
    zeta = fdtool.visco_constq_zeta(NM, fq_min, fq_max, ts)
    if benchmark:
        zeta = F32(0.0)
    taup = (F32(NM) * zeta / taup).astype(F32)
    taus = (F32(NM) * zeta / taus).astype(F32)
    # relaxed_medium with visco_chi (src/swpc_psv/m_medium.f90:199-200)
    rho_beta2 = mu.copy()
    rho_alpha2 = (lam + 2 * mu).astype(F32)
    chi = np.vectorize(lambda t: fdtool.visco_chi(NM, ts, t, fq_ref), otypes=[F32])
    ut, inv = np.unique(np.concatenate([taus.ravel(), taup.ravel()]), return_inverse=True)
    cv = chi(ut)
    chis = cv[inv[:taus.size]].reshape(taus.shape)
    chip = cv[inv[taus.size:]].reshape(taup.shape)
    mu = (rho_beta2 / chis ** 2).astype(F32)
    lam = (rho_alpha2 / chip ** 2 - 2 * mu).astype(F32)

    kfs = np.zeros(nx + 6, np.int32)
    kob = np.zeros(nx + 6, np.int32)
    f, ob = _detect(lam[2:nx + 5], mu[2:nx + 5], True, 1, nz, 0)
    kfs[2:nx + 5] = f
    kob[2:nx + 5] = ob
    from numpy.lib.stride_tricks import sliding_window_view as swv
    wf = swv(kfs, 6)[1:nx + 1]
    wo = swv(kob, 6)[1:nx + 1]
    kfs_top = np.zeros_like(kfs)
    kfs_bot = np.zeros_like(kfs)
    kob_top = np.zeros_like(kfs)
    kob_bot = np.zeros_like(kfs)
    kfs_top[3:nx + 3] = np.maximum(wf.min(axis=1) - 2, 1)
    kfs_bot[3:nx + 3] = np.minimum(wf.max(axis=1) + 2, nz)
    kob_top[3:nx + 3] = np.maximum(wo.min(axis=1) - 2, 1)
    kob_bot[3:nx + 3] = np.minimum(wo.max(axis=1) + 2, nz)
    vmin, vmax = _minmax(lam[3:nx + 3, 3:nz + 3][None], mu[3:nx + 3, 3:nz + 3][None],
                         rho[3:nx + 3, 3:nz + 3][None], kfs[3:nx + 3][None])
    return Medium2D(rho, lam, mu, taup, taus, kfs, kob, kfs_top, kfs_bot, kob_top, kob_bot, ts, float(zeta), vmin, vmax)


# ---------------------------------------------------------------------------
# model builders
# ---------------------------------------------------------------------------
def vmax_clip3(g: Grid3D):
    """vmax of the random-media clamp: 6/7 dh / dt with dh evaluated in
    double precision and stored in real(SP) (m_vmodel_lhm_rmed.f90:92-94)."""
    dh = F32(1.0 / math.sqrt(1.0 / g.dx ** 2 + 1.0 / g.dy ** 2 + 1.0 / g.dz ** 2))
    return F32(F32(6.0) / F32(7.0)) * dh / F32(g.dt)


def vmax_clip2(g: Grid2D):
    dh = F32(1.0 / math.sqrt(1.0 / g.dx ** 2 + 1.0 / g.dz ** 2))
    return F32(F32(6.0) / F32(7.0)) * dh / F32(g.dt)


def layered3(g: Grid3D, layers, vcut=0.0, xi=None, rhomin=1.0, munk=False):
    """rho, lam, mu, qp, qs over the memory box for a 1-D layered model,
    optionally perturbed by a random field xi of the box's shape."""
    zc = g.zc
    if xi is None:
        cols = vmodel_lhm(zc, layers, vcut, munk=munk)
        return tuple(np.ascontiguousarray(np.broadcast_to(c, g.shape)).astype(F32) for c in cols)
    return vmodel_lhm(zc, layers, vcut, xi=xi, vmax_clip=vmax_clip3(g), rhomin=rhomin, munk=munk)


def layered2(g: Grid2D, layers, vcut=0.0, xi=None, rhomin=1.0, munk=False):
    zc = g.zc
    if xi is None:
        cols = vmodel_lhm(zc, layers, vcut, munk=munk)
        return tuple(np.ascontiguousarray(np.broadcast_to(c, g.shape)).astype(F32) for c in cols)
    return vmodel_lhm(zc, layers, vcut, xi=xi, vmax_clip=vmax_clip2(g), rhomin=rhomin, munk=munk)


def uniform(shape, zc, vp=5.0, vs=None, rho=2.7, qp=1e6, qs=1e6, topo=0.0):
    """vmodel_uni: homogeneous half-space below z = topo, air above."""
    if vs is None:
        vs = vp / math.sqrt(3.0)
    return layered_from_list(shape, zc, [Layer(topo, rho, vp, vs, qp, qs)])


def layered_from_list(shape, zc, layers):
    cols = vmodel_lhm(zc, layers, 0.0)
    return tuple(np.ascontiguousarray(np.broadcast_to(c, shape)).astype(F32) for c in cols)
