"""SAC output (src/shared/m_sac.f90 sac__whdr/wsac_s; src/swpc_3d/m_wav.f90
wav__write): header bytes against files written by the reference's own
sac__write (golden fixtures and, where built, live calls), read-back
round trips and the wav/<title>.3d.<stnm>.<cmp>.sac naming.
"""
import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

from openswpc_amd import sac
from oracle.oracle import ref_lib

G = json.loads((Path(__file__).parent / "golden" / "ref_sac.json").read_text())


def header_of(case):
    h = sac.SacHeader()
    for k, v in zip(G["fh_order"], case["fh"]):
        setattr(h, k, v)
    for k, v in zip(G["ih_order"], case["ih"]):
        setattr(h, k, v)
    # Fortran character assignment truncates to the field length (8)
    h.kstnm = case["kstnm"][:8]
    h.kcmpnm = case["kcmpnm"][:8]
    return h


@pytest.mark.parametrize("n", range(len(G["cases"])))
def test_sac_bytes_match_reference_writer(n, tmp_path):
    case = G["cases"][n]
    p = tmp_path / "x.sac"
    sac.write_sac(p, header_of(case), np.array(case["data"], np.float32))
    assert p.read_bytes().hex() == case["bytes"]


def test_sac_delta_truncation():
    """delta is stored as real(int(delta*1d7))/1e7: truncated, not rounded."""
    h = sac.SacHeader()
    h.npts = 0
    for d, want in [(0.3, np.float32(0.3)), (0.0123456789, np.float32(123456) / np.float32(1e7)), (0.05, np.float32(0.05)),
                    (0.12345678, np.float32(1234567) / np.float32(1e7))]:
        h.delta = d
        got = np.frombuffer(h.pack()[:4], np.float32)[0]
        assert got == want, (d, got, want)


def test_sac_defaults_and_roundtrip(tmp_path):
    h = sac.SacHeader()
    raw = h.pack()
    assert len(raw) == 632
    f = np.frombuffer(raw[:280], np.float32)
    i = np.frombuffer(raw[280:420], np.int32)
    lg = np.frombuffer(raw[420:440], np.int32)
    assert np.all(f == -12345.0)
    assert i[77 - 71] == 6 and i[86 - 71] == 1 and i[80 - 71] == -12345
    assert list(lg) == [1, 0, 1, 0, 0]
    assert raw[440:448] == b"-12345  "
    h.npts = 5
    h.kstnm = "AB12"
    h.user3 = 0.25
    h.b = -0.5
    d = np.linspace(-1, 1, 5, dtype=np.float32)
    sac.write_sac(tmp_path / "r.sac", h, d)
    h2, d2 = sac.read_sac(tmp_path / "r.sac")
    assert h2.kstnm == "AB12" and h2.npts == 5 and h2.user3 == 0.25 and h2.b == -0.5
    assert np.array_equal(d2, d)
    with pytest.raises(ValueError):
        h.npts = 9
        sac.write_sac(tmp_path / "bad.sac", h, d)


def test_live_reference_writer(tmp_path):
    L = ref_lib()
    if L is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(11)
    for trial in range(10):
        npts = int(rng.integers(1, 200))
        fh = [float(v) for v in rng.uniform(-100, 100, 11)]
        fh[0] = float(rng.uniform(1e-4, 0.5))
        ih = [npts, int(rng.integers(5, 8)), int(rng.integers(1900, 2100)), int(rng.integers(1, 366))]
        dat = rng.standard_normal(npts).astype(np.float32)
        fn = str(tmp_path / f"ref{trial}.sac").encode()
        L.ref_wsac(fn, len(fn), (C.c_double * 11)(*fh), (C.c_int * 4)(*ih), b"ST%02d    " % trial, b"Vz      ",
                   npts, dat.ctypes.data_as(C.POINTER(C.c_float)))
        case = {"fh": fh, "ih": ih, "kstnm": f"ST{trial:02d}", "kcmpnm": "Vz"}
        p = tmp_path / f"mine{trial}.sac"
        sac.write_sac(p, header_of(case), dat)
        assert p.read_bytes() == Path(fn.decode()).read_bytes(), trial


def test_write_wav_layout(tmp_path):
    """One file per station and component, named as export_wav__sac does,
    with the header fields initialize_sac_header sets."""
    from openswpc_amd.source import Station
    st = [Station(1.0, 2.0, 0.0, "N001", "fsb"), Station(-3.0, 4.0, 0.5, "N002", "dep")]
    ntw = 12
    meta = sac.WavMeta(title="swpc", dt=0.01, ntdec_w=5, ntw=ntw, tbeg=0.0, evdp=10.0, sx0=0.0, sy0=0.0,
                       mag=5.0, otim=0.1, m6=(0.1, 0.2, 0.3, 0.4, 0.5, 0.6), clon=139.0, clat=35.0, phi=0.0,
                       exedate=0)
    vel = np.random.default_rng(0).standard_normal((ntw, 2, 3)).astype(np.float32)
    disp = np.random.default_rng(1).standard_normal((ntw, 2, 3))
    paths = sac.write_wav(tmp_path, meta, st, vel=vel, disp=disp)
    names = sorted(Path(p).name for p in paths)
    assert names == sorted(f"swpc.3d.{s}.{c}.sac" for s in ("N001", "N002") for c in ("Vx", "Vy", "Vz", "Ux", "Uy", "Uz"))
    h, d = sac.read_sac(tmp_path / "wav" / "swpc.3d.N002.Vy.sac")
    assert np.array_equal(d, vel[:, 1, 1])
    assert h.idep == 7 and h.kcmpnm == "Vy" and h.cmpaz == 90.0 and h.cmpinc == 90.0
    assert h.user5 == np.float32(0.6) and h.user6 == 139.0 and h.o == np.float32(0.1)
    assert h.stdp == 500.0 and h.delta == np.float32(0.05) and h.npts == ntw
    assert abs(h.dist - 5.0) < 1e-6 and abs(h.az - np.degrees(np.arctan2(4, -3))) < 1e-4
    h, d = sac.read_sac(tmp_path / "wav" / "swpc.3d.N001.Uz.sac")
    assert h.idep == 6 and np.array_equal(d, disp[:, 0, 2].astype(np.float32))
    # P-SV: two components, in-plane moment entries only
    pv = sac.write_wav(tmp_path, meta, st, vel=vel[:, :, :2], dim="psv")
    assert sorted(Path(p).name for p in pv) == sorted(f"swpc.psv.{s}.{c}.sac" for s in ("N001", "N002") for c in ("Vx", "Vz"))
    h, _ = sac.read_sac(tmp_path / "wav" / "swpc.psv.N002.Vz.sac")
    assert h.user0 == np.float32(0.1) and h.user1 == -12345.0 and h.user4 == np.float32(0.5)
    assert h.dist == 3.0 and h.az == 180.0
