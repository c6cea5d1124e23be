"""Writes tests/golden/ref_fdtool.json from the reference's own Fortran
modules (src/shared/m_fdtool.f90, m_sac.f90), compiled unmodified by
oracle/Makefile into oracle/_ref/libswpcref.so.

Run in the build container (where /root/reference exists):
    make -C oracle ref && python tests/golden/make_ref_vectors.py
The JSON holds inputs and the reference's outputs only.
"""
import ctypes as C
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle.oracle import ref_lib  # noqa: E402

F = C.c_float


def main():
    L = ref_lib()
    if L is None:
        raise SystemExit("oracle/_ref/libswpcref.so not built")
    L.ref_momentrate.restype = F
    L.ref_ikupper.restype = F
    L.ref_seismic_moment.restype = F
    L.ref_visco_constq_zeta.restype = F
    L.ref_visco_chi.restype = F
    L.ref_x2i.restype = C.c_int
    L.ref_i2x.restype = F
    rng = np.random.default_rng(2024)
    out = {}

    # sdr2moment over random and special mechanisms
    sdr = [(0, 90, 0), (0, 90, 180), (45, 45, 90), (30, 60, -90), (360, 0, 0), (-360, -90, -180)]
    sdr += [tuple(float(v) for v in (rng.uniform(-360, 360), rng.uniform(-90, 90), rng.uniform(-180, 180)))
            for _ in range(40)]
    res = []
    for s, d, r in sdr:
        m = (F * 6)()
        L.ref_sdr2moment(F(s), F(d), F(r), m)
        res.append([float(v) for v in m])
    out["sdr2moment"] = {"in": [list(map(float, x)) for x in sdr], "out": res}

    # source time functions at many t (kind: 0 boxcar .. 5 texp)
    ts, tr = 0.3, 1.7
    tt = [float(v) for v in np.concatenate([np.linspace(-0.5, 3.0, 141), [0.3, 1.15, 2.0, 0.725, 1.575]])]
    out["momentrate"] = {"ts": ts, "tr": tr, "t": tt,
                         "out": {str(k): [float(L.ref_momentrate(F(t), k, F(ts), F(tr))) for t in tt] for k in range(6)}}
    out["ikupper"] = {"ts": ts, "tr": tr, "t": tt, "out": [float(L.ref_ikupper(F(t), F(ts), F(tr))) for t in tt]}
    mws = [float(v) for v in np.linspace(-1, 9, 21)]
    out["seismic_moment"] = {"mw": mws, "out": [float(L.ref_seismic_moment(F(m))) for m in mws]}

    # viscoelastic constants
    vis = []
    for nm, fmin, fmax in [(3, 0.05, 5.0), (3, 0.1, 10.0), (1, 0.05, 5.0), (2, 0.02, 2.0), (5, 0.01, 20.0)]:
        ts_ = (F * nm)()
        L.ref_visco_set_relaxtime(nm, F(fmin), F(fmax), ts_)
        z = float(L.ref_visco_constq_zeta(nm, F(fmin), F(fmax), ts_))
        chis = [(tau, f, float(L.ref_visco_chi(nm, ts_, F(tau), F(f))))
                for tau in (0.0, 0.01, 0.05, 0.3) for f in (0.5, 1.0, 3.0)]
        vis.append({"nm": nm, "fmin": fmin, "fmax": fmax, "ts": [float(v) for v in ts_], "zeta": z, "chi": chis})
    out["visco"] = vis

    # index maps (ceiling in single precision, including exact cell edges)
    maps = []
    for xbeg, dx in [(-64.0, 0.25), (-40.0, 0.1), (-7.5, 0.25), (-2.0, 0.5)]:
        xs = [xbeg + dx * v for v in (0.0, 0.5, 1.0, 7.0, 7.25, 100.0, 100.999)] + \
             [float(v) for v in rng.uniform(xbeg, xbeg + 100 * dx, 12)]
        maps.append({"xbeg": xbeg, "dx": dx, "x": xs, "i": [int(L.ref_x2i(F(x), F(xbeg), F(dx))) for x in xs],
                     "i2x": [float(L.ref_i2x(i, F(xbeg), F(dx))) for i in range(-2, 12)]})
    out["maps"] = maps

    # vcheck clamp
    vc = []
    for vp, vs, rho, xi in [(6.0, 3.5, 2.7, 0.05), (13.0, 7.0, 3.0, 0.2), (1.0, 0.4, 1.5, -0.3), (5.0, 3.0, 0.5, -0.4)]:
        a, b, c = F(vp), F(vs), F(rho)
        fl = (C.c_int * 3)()
        L.ref_vcheck(C.byref(a), C.byref(b), C.byref(c), F(xi), F(0.5), F(12.0), F(1.0), fl)
        vc.append({"in": [vp, vs, rho, xi], "vmin": 0.5, "vmax": 12.0, "rhomin": 1.0,
                   "out": [a.value, b.value, c.value], "flags": list(fl)})
    out["vcheck"] = vc

    p = Path(__file__).with_name("ref_fdtool.json")
    p.write_text(json.dumps(out, indent=1))
    print("wrote", p)
    sac_vectors(L, rng)
    readini_vectors(L)


def sac_vectors(L, rng):
    """SAC files written by sac__init + sac__write (m_sac.f90) for a few
    header/data sets, stored as hex bytes."""
    import tempfile
    cases = []
    deltas = [0.05, 0.01, 0.003, 0.1 * 3, 0.0123456789, 1.0 / 3.0]
    for n, delta in enumerate(deltas):
        npts = [1, 7, 16, 33, 5, 64][n]
        fh = [delta, -1.5 + n, 20.0 + n, -3.25, 4.5, 139.7, 35.6, 140.1, 36.2, 10.5 + n, 1 / np.sqrt(3.0)]
        ih = [npts, [7, 6, 5][n % 3], 2025 + n, 1 + 60 * n]
        kst = ["ST01", "N.KI2H", "ABCDEFGHIJ", "", "X", "stn-9"][n]
        kcm = ["Vx", "Uz", "Sxy", "Vy", "Exz", "Vz"][n]
        dat = (rng.standard_normal(npts) * 10.0 ** (n - 3)).astype(np.float32)
        with tempfile.TemporaryDirectory() as td:
            fn = (Path(td) / "t.sac").as_posix().encode()
            a = kst.encode()[:8].ljust(8)
            b = kcm.encode()[:8].ljust(8)
            L.ref_wsac(fn, len(fn), (C.c_double * 11)(*fh), (C.c_int * 4)(*ih), a, b, npts,
                       dat.ctypes.data_as(C.POINTER(C.c_float)))
            raw = Path(fn.decode()).read_bytes()
        cases.append({"fh": fh, "ih": ih, "kstnm": kst, "kcmpnm": kcm, "data": [float(v) for v in dat],
                      "bytes": raw.hex()})
    p = Path(__file__).with_name("ref_sac.json")
    p.write_text(json.dumps({"fh_order": ["delta", "b", "e", "depmin", "depmax", "stlo", "stla", "evlo", "evla",
                                          "evdp", "user0"],
                             "ih_order": ["npts", "idep", "nzyear", "nzjday"], "cases": cases}, indent=1))
    print("wrote", p)



INI_TEXT = """# parameter file exercising readini's parsing
  ! indented comment
title = 'swpc test'   ! quoted value with a blank
odir = ./out/dir      # unquoted: the item ends at the slash
nx= 512, 33
nxs   =  9
dx = 1.5d-1
dt = 5e-3
fq = .5
big = 1.0E+02
neg = -12
ndig  =  007
fl1 = .true.
fl2 = T
fl3 = false
fl4 = .F.
fl5 = .t.
str2 = "dq ""x"" y"
str3 = a,b
str4 = ${SWPC_TEST_DIR}/x${SWPC_TEST_USER}y
str5 = pre{X}post
str6 = '${SWPC_TEST_DIR} and more'
tab\t= 7
nx = 1000
 sp_key    =    3.25
#nz = 99
nz = 64
dx2 = 2.
"""
INI_ENV = {"SWPC_TEST_DIR": "/data/swpc", "SWPC_TEST_USER": "me"}
INI_QUERIES = [("title", 3), ("odir", 3), ("str2", 3), ("str3", 3), ("str4", 3), ("str5", 3), ("str6", 3),
               ("missing", 3), ("nx", 0), ("nxs", 0), ("neg", 0), ("ndig", 0), ("nz", 0), ("tab", 0),
               ("missing", 0), ("dx", 1), ("dx", 2), ("dt", 1), ("dt", 2), ("fq", 1), ("big", 2), ("sp_key", 1),
               ("sp_key", 2), ("dx2", 2), ("missing", 2), ("fl1", 4), ("fl2", 4), ("fl3", 4), ("fl4", 4),
               ("fl5", 4), ("missing", 4), ("n", 0), ("title ", 3)]


def readini_vectors(L):
    """Values returned by readini (m_readini.f90) for INI_TEXT."""
    import os
    import tempfile
    os.environ.update(INI_ENV)
    out = []
    with tempfile.TemporaryDirectory() as td:
        fn = (Path(td) / "t.inf").as_posix()
        Path(fn).write_text(INI_TEXT)
        fb = fn.encode()
        for key, kind in INI_QUERIES:
            i, r, d, l = C.c_int(-7), C.c_float(2.5), C.c_double(2.5), C.c_int(1)
            s = C.create_string_buffer(b"dflt", 256)
            kb = key.encode()
            L.ref_readini(fb, len(fb), kb, len(kb), kind, C.byref(i), C.byref(r), C.byref(d), s, C.byref(l))
            val = [i.value, r.value, d.value, s.value.decode().rstrip(" "), bool(l.value)][kind]
            out.append({"key": key, "kind": kind, "value": val})
    p = Path(__file__).with_name("ref_readini.json")
    p.write_text(json.dumps({"text": INI_TEXT, "env": INI_ENV, "defaults": {"i": -7, "r": 2.5, "c": "dflt", "l": True},
                             "queries": out}, indent=1))
    print("wrote", p)


if __name__ == "__main__":
    main()
