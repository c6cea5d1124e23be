"""Geographic <-> model Cartesian coordinates.

The model frame is a transverse Mercator (Gauss-Krueger) projection of the
GRS80 ellipsoid about the centre (clon, clat) with scale factor 0.9999,
rotated clockwise by the azimuth phi (for phi = 0: x north, y east), in km.
This is what src/shared/m_geomap.f90:19-64 (geomap__g2c / geomap__c2g) does
on top of src/shared/m_gk.f90:40-230.

The projection follows the published method of the Geospatial Information
Authority of Japan (K. Kawase, "A general formula for calculating meridian
arc length and its application to coordinate conversion in the
Gauss-Krueger projection", Bull. GSI 59, 2011): Krueger's series in the
third flattening n, to n^5 for the forward and inverse series and n^6 for
the conformal-to-geodetic latitude series.  The series coefficients are
the mathematical expansions of that method, so they are the ones the
reference uses.

Single-precision inputs are widened to float64, projected in float64 and
rounded back to float32, as the reference's SP wrappers do
(m_gk.f90:89-103, 158-172); the rotation runs in float32 with the angle
converted as ``real(pi/180 * phi)`` (src/shared/m_std.f90:132-139).
"""
from __future__ import annotations

import math

import numpy as np

F32 = np.float32

A_GRS80 = 6378137.0          # equatorial radius [m]
F_INV = 298.257222101        # inverse flattening of GRS80
SCALE = 0.9999               # central-meridian scale factor


def _coefficients():
    n = 1.0 / (2.0 * F_INV - 1.0)
    # forward Krueger series (geodetic -> projected)
    alpha = [
        n / 2 - 2 * n**2 / 3 + 5 * n**3 / 16 + 41 * n**4 / 180 - 127 * n**5 / 288,
        13 * n**2 / 48 - 3 * n**3 / 5 + 557 * n**4 / 1440 + 281 * n**5 / 630,
        61 * n**3 / 240 - 103 * n**4 / 140 + 15061 * n**5 / 26880,
        49561 * n**4 / 161280 - 179 * n**5 / 168,
        34729 * n**5 / 80640,
    ]
    # inverse series (projected -> conformal sphere)
    beta = [
        n / 2 - 2 * n**2 / 3 + 37 * n**3 / 96 - n**4 / 360 - 81 * n**5 / 512,
        n**2 / 48 + n**3 / 15 - 437 * n**4 / 1440 + 46 * n**5 / 105,
        17 * n**3 / 480 - 37 * n**4 / 840 - 209 * n**5 / 4480,
        4397 * n**4 / 161280 - 11 * n**5 / 504,
        4583 * n**5 / 161280,
    ]
    # conformal latitude -> geodetic latitude
    delta = [
        2 * n - 2 * n**2 / 3 - 2 * n**3 + 116 * n**4 / 45 + 26 * n**5 / 45 - 2854 * n**6 / 675,
        7 * n**2 / 3 - 8 * n**3 / 5 - 227 * n**4 / 45 + 2704 * n**5 / 315 + 2323 * n**6 / 945,
        56 * n**3 / 15 - 136 * n**4 / 35 - 1262 * n**5 / 105 + 73814 * n**6 / 2835,
        4279 * n**4 / 630 - 332 * n**5 / 35 - 399572 * n**6 / 14175,
        4174 * n**5 / 315 - 144838 * n**6 / 6237,
        601676 * n**6 / 22275,
    ]
    # meridian arc S(phi) = SCALE a / (1+n) (A0 phi + sum_j Aj sin 2j phi)
    arc = [
        1 + n**2 / 4 + n**4 / 64,
        -1.5 * (n - n**3 / 8 - n**5 / 64),
        15 / 16 * (n**2 - n**4 / 4),
        -35 / 48 * (n**3 - 5 * n**5 / 16),
        315 / 512 * n**4,
        -693 / 1280 * n**5,
    ]
    return n, alpha, beta, delta, arc


N3, ALPHA, BETA, DELTA, ARC = _coefficients()
ECC = 2.0 * math.sqrt(N3) / (1.0 + N3)     # first eccentricity
ABAR = SCALE * A_GRS80 / (1.0 + N3) * ARC[0]


def meridian_arc(lat_rad):
    """Scaled meridian arc length [m] from the equator to lat_rad."""
    s = ARC[0] * lat_rad
    for j in range(1, 6):
        s = s + ARC[j] * np.sin(2 * j * lat_rad)
    return s * (SCALE * A_GRS80 / (1.0 + N3))


def _atanh(x):
    return 0.5 * np.log((1.0 + x) / (1.0 - x))


def gk_ll2xy(lon, lat, lon0, lat0):
    """Geodetic (deg) -> projected (km, x north, y east), float64 throughout."""
    lam = np.radians(np.asarray(lon, np.float64))
    phi = np.radians(np.asarray(lat, np.float64))
    lam0 = math.radians(float(lon0))
    phi0 = math.radians(float(lat0))
    dl = lam - lam0
    sphi = np.sin(phi)
    t = np.sinh(_atanh(sphi) - ECC * np.arctanh(ECC * sphi))   # tan of the conformal latitude
    xi = np.arctan(t / np.cos(dl))
    eta = _atanh(np.sin(dl) / np.sqrt(1.0 + t * t))
    x, y = xi, eta
    for j in range(1, 6):
        x = x + ALPHA[j - 1] * np.sin(2 * j * xi) * np.cosh(2 * j * eta)
        y = y + ALPHA[j - 1] * np.cos(2 * j * xi) * np.sinh(2 * j * eta)
    return (ABAR * x - meridian_arc(phi0)) / 1000.0, ABAR * y / 1000.0


def gk_xy2ll(x, y, lon0, lat0):
    """Projected (km) -> geodetic (deg), float64 throughout."""
    lam0 = math.radians(float(lon0))
    phi0 = math.radians(float(lat0))
    xi = (np.asarray(x, np.float64) * 1000.0 + meridian_arc(phi0)) / ABAR
    eta = np.asarray(y, np.float64) * 1000.0 / ABAR
    xi2, eta2 = xi, eta
    for j in range(1, 6):
        xi2 = xi2 - BETA[j - 1] * np.sin(2 * j * xi) * np.cosh(2 * j * eta)
        eta2 = eta2 - BETA[j - 1] * np.cos(2 * j * xi) * np.sinh(2 * j * eta)
    chi = np.arcsin(np.sin(xi2) / np.cosh(eta2))
    lam = lam0 + np.arctan(np.sinh(eta2) / np.cos(xi2))
    phi = chi
    for j in range(1, 7):
        phi = phi + DELTA[j - 1] * np.sin(2 * j * chi)
    return np.degrees(lam), np.degrees(phi)


def _rot(phi):
    r = F32(math.pi / 180.0 * float(F32(phi)))
    return F32(np.cos(r)), F32(np.sin(r))


def g2c(lon, lat, clon, clat, phi):
    """geomap__g2c: (lon, lat) in degrees -> model (x, y) in km (float32)."""
    xx, yy = gk_ll2xy(np.asarray(lon, F32).astype(np.float64), np.asarray(lat, F32).astype(np.float64),
                      float(F32(clon)), float(F32(clat)))
    xx, yy = np.asarray(xx).astype(F32), np.asarray(yy).astype(F32)
    c, s = _rot(phi)
    x = (c * xx + s * yy).astype(F32)
    y = (-s * xx + c * yy).astype(F32)
    return (F32(x), F32(y)) if x.ndim == 0 else (x, y)


def c2g(x, y, clon, clat, phi):
    """geomap__c2g: model (x, y) in km -> (lon, lat) in degrees (float32)."""
    c, s = _rot(phi)
    x, y = np.asarray(x, F32), np.asarray(y, F32)
    xx = (c * x - s * y).astype(F32)
    yy = (s * x + c * y).astype(F32)
    lon, lat = gk_xy2ll(xx.astype(np.float64), yy.astype(np.float64), float(F32(clon)), float(F32(clat)))
    return F32(lon) if np.ndim(lon) == 0 else lon.astype(F32), F32(lat) if np.ndim(lat) == 0 else lat.astype(F32)
