"""Random-media synthesis (von Karman / Gaussian / exponential PSDF).

Follows the method of ``rmedia__3dgen_ani`` / ``rmedia__2dgen_ani``
(src/shared/m_rmedia.f90:40-173): uniform noise -> FFT -> spectral
whitening -> multiply by sqrt(PSDF) -> zero DC -> inverse FFT -> demean ->
scale to rms epsilon.  The noise comes from numpy's PCG64 seeded stream, so
individual realisations differ from the reference's gfortran RNG while the
statistics (PSDF, rms) are the same.  The volume is periodic and is tiled
over the model like ``rdrmed__3d`` does (src/shared/m_rdrmed.f90:111-132).
"""
from __future__ import annotations

import math

import numpy as np


def psdf_vonkarman(m2a2, a3, epsil, kappa, ndim=3):
    """von Karman PSDF (m_rmedia.f90:716-732 for 3-D) with m2a2 = sum (k a)^2."""
    if ndim == 3:
        c = 8 * math.pi ** 1.5 * epsil ** 2 * a3 * math.gamma(kappa + 1.5) / math.gamma(kappa)
        return c / (1 + m2a2) ** (kappa + 1.5)
    c = 4 * math.pi * kappa * epsil ** 2 * a3
    return c / (1 + m2a2) ** (kappa + 1.0)


def psdf_gauss(m2a2, a3, epsil, ndim=3):
    return epsil ** 2 * a3 * (math.sqrt(math.pi) ** ndim) * np.exp(-m2a2 / 4)


def psdf_exp(m2a2, a3, epsil, ndim=3):
    if ndim == 3:
        return 8 * math.pi * epsil ** 2 * a3 / (1 + m2a2) ** 2
    return 2 * math.pi * epsil ** 2 * a3 / (1 + m2a2) ** 1.5


PTYPES = ("vonkarman", "gauss", "exp")


def generate(shape, d, a, epsil, kappa=0.5, ptype="vonkarman", seed=0, device=False):
    """Periodic random field of the given shape (nx, ny[, nz]) with grid
    widths d and correlation lengths a (per axis), rms epsil; float32.

    ``device=True`` runs the transforms, the filter and the scaling on the
    GPU (swpc_rmedia_generate: hipFFT in FP32, the PSDF and the reductions
    in FP64) on the same noise; it raises if the HIP library is absent."""
    shape = tuple(int(n) for n in shape)
    nd = len(shape)
    d = tuple(float(x) for x in (d if np.iterable(d) else [d] * nd))
    a = tuple(float(x) for x in (a if np.iterable(a) else [a] * nd))
    rng = np.random.default_rng(seed)
    noise = rng.random(shape)
    if device:
        return _generate_device(shape, d, a, epsil, kappa, ptype, noise)
    B = np.fft.rfftn(noise)
    mag = np.abs(B)
    mag[mag == 0] = 1.0
    B /= mag
    ks = [2 * np.pi * np.fft.fftfreq(n, dd) for n, dd in zip(shape[:-1], d[:-1])]
    ks.append(2 * np.pi * np.fft.rfftfreq(shape[-1], d[-1]))
    m2a2 = 0.0
    for ax, (k, aa) in enumerate(zip(ks, a)):
        sh = [1] * nd
        sh[ax] = -1
        m2a2 = m2a2 + (k.reshape(sh) * aa) ** 2
    a3 = float(np.prod(a))
    if ptype == "vonkarman":
        p = psdf_vonkarman(m2a2, a3, epsil, kappa, nd)
    elif ptype == "gauss":
        p = psdf_gauss(m2a2, a3, epsil, nd)
    elif ptype == "exp":
        p = psdf_exp(m2a2, a3, epsil, nd)
    else:
        raise ValueError(ptype)
    B *= np.sqrt(p * float(np.prod(shape)) * float(np.prod(d)))
    B[(0,) * nd] = 0.0
    f = np.fft.irfftn(B, s=shape, axes=tuple(range(nd)))
    f -= f.mean()
    rms = math.sqrt(float(np.mean(f * f)))
    f *= epsil / rms
    return f.astype(np.float32)


def _generate_device(shape, d, a, epsil, kappa, ptype, noise):
    import ctypes as C

    from ._lib import PD, PF, PI_, check, lib, ptr
    if ptype not in PTYPES:
        raise ValueError(ptype)
    nd = len(shape)
    sh = np.asarray(shape, np.int32)
    dd, aa = np.asarray(d, np.float64), np.asarray(a, np.float64)
    nz = np.ascontiguousarray(noise, np.float32)
    out = np.empty(shape, np.float32)
    check(lib().swpc_rmedia_generate(nd, ptr(sh, C.c_int), ptr(dd, C.c_double), ptr(aa, C.c_double), float(epsil),
                                     float(kappa), PTYPES.index(ptype), ptr(nz, C.c_float), ptr(out, C.c_float)),
          "swpc_rmedia_generate")
    return out
