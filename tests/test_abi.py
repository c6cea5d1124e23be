"""The C ABI without a GPU: libswpc_hip.so loads, exports every function
that include/swpc_hip.h declares, every one has a ctypes prototype, and
argument validation answers with error codes and messages before any
device call.
"""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "swpc_hip.h"


def declared():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(swpc\w+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def L():
    from openswpc_amd import _lib
    if not _lib.LIB_PATH.exists():
        from openswpc_amd.build import build_library
        build_library()
    return _lib.lib()


def test_header_declares_the_boundary():
    names = declared()
    for must in ("swpc3d_create", "swpc3d_update_stress", "swpc3d_update_vel", "swpc3d_stressglut",
                 "swpc3d_store_receivers", "swpc3d_vmax", "swpc2d_create", "swpc2d_step", "swpc_last_error"):
        assert must in names
    assert len(names) >= 45


def test_library_exports_every_declared_symbol(L):
    from openswpc_amd import _lib
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    # every declared entry point has a ctypes prototype in the host wrapper
    assert not [n for n in declared() if n not in _lib._PROTOS]


def test_validation_without_device(L):
    from openswpc_amd._lib import Desc3, Desc2
    assert L.swpc_version() >= 10000
    h = C.c_void_p()
    assert L.swpc3d_create(None, C.byref(h)) == -1
    assert b"null" in L.swpc_last_error()
    d = Desc3(nx=4, ny=64, nz=64, na=8, nm=3, precision=0, dx=0.1, dy=0.1, dz=0.1, dt=0.01)
    assert L.swpc3d_create(C.byref(d), C.byref(h)) == -1
    assert b"too small" in L.swpc_last_error()
    d = Desc3(nx=64, ny=64, nz=64, na=8, nm=2, precision=0, dx=0.1, dy=0.1, dz=0.1, dt=0.01)
    assert L.swpc3d_create(C.byref(d), C.byref(h)) == -1
    assert b"NM = 3" in L.swpc_last_error()
    d = Desc3(nx=64, ny=64, nz=64, na=8, nm=3, precision=7, dx=0.1, dy=0.1, dz=0.1, dt=0.01)
    assert L.swpc3d_create(C.byref(d), C.byref(h)) == -1
    d = Desc3(nx=64, ny=64, nz=64, na=40, nm=3, precision=0, dx=0.1, dy=0.1, dz=0.1, dt=0.01)
    assert L.swpc3d_create(C.byref(d), C.byref(h)) == -1
    assert b"absorber" in L.swpc_last_error()
    d2 = Desc2(nx=64, nz=64, na=8, nm=3, precision=0, dx=0.1, dz=-0.1, dt=0.01)
    assert L.swpc2d_create(C.byref(d2), C.byref(h)) == -1
    # calls on a null context are rejected, not dereferenced
    assert L.swpc3d_update_stress(None) != 0
    assert L.swpc2d_update_vel(None) != 0
