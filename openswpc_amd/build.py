"""Build recipe for the HIP library (gfx950) and the CPU oracle.

``build_library()`` compiles every ``csrc/*.hip`` with hipcc for gfx950 and
links ``openswpc_amd/libswpc_hip.so`` in-tree.  ``build_fortran_host()``
compiles the Fortran bindings (``include/swpc_hip.f90``) and the Fortran
time-loop host (``fhost/``) with amdflang into ``fhost/swpc3d_fdrive``.  ``build_oracle()`` runs the
Makefile under ``oracle/`` (the C restatement used by the tests, and, when
the reference sources are present, the amdflang build of the reference's
shared Fortran modules into ``oracle/_ref``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libswpc_hip.so"
BUILD_DIR = ROOT / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SWPC_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: keep hipcc from forming packed FP32 instructions.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(map(str, cmd))}\n{out}")
    return out


def build_library(verbose: bool = False) -> Path:
    """Compile csrc/*.hip for gfx950 and link libswpc_hip.so in-tree."""
    BUILD_DIR.mkdir(exist_ok=True)
    objs = []
    for src in sorted(CSRC.glob("*.hip")):
        obj = BUILD_DIR / (src.stem + ".o")
        out = _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
        if verbose and out.strip():
            print(out)
        objs.append(obj)
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", str(tmp), *map(str, objs), "-lhipfft"])
    os.replace(tmp, LIB)
    return LIB


FLANG = os.environ.get("SWPC_FLANG", "/opt/rocm/lib/llvm/bin/flang")
FDRIVE = PKG / "fhost" / "swpc3d_fdrive"
FDRIVE2 = PKG / "fhost" / "swpc2d_fdrive"


def build_fortran_host(verbose: bool = False):
    """Fortran bindings + the 3-D and P-SV drop-in hosts; linked against
    libswpc_hip.so with an rpath to the package directory.  Returns the 3-D
    executable, or None where no Fortran compiler is installed."""
    if not Path(FLANG).exists():
        return None
    mod = BUILD_DIR / "fmod"
    mod.mkdir(parents=True, exist_ok=True)

    def obj(src):
        o = mod / (src.stem + ".o")
        _run([FLANG, "-O2", "-fPIC", "-module-dir", str(mod), "-I", str(mod), "-c", str(src), "-o", str(o)])
        return o

    fh = PKG / "fhost"
    binding = obj(ROOT / "include" / "swpc_hip.f90")
    for exe, srcs in ((FDRIVE, ("m_kernel_hip.f90", "swpc3d_fdrive.f90")),
                      (FDRIVE2, ("m_kernel_psv_hip.f90", "swpc2d_fdrive.f90"))):
        objs = [binding] + [obj(fh / s) for s in srcs]
        tmp = exe.with_name(exe.name + ".tmp")
        _run([FLANG, "-o", str(tmp), *map(str, objs), f"-L{PKG}", "-lswpc_hip", "-Wl,-rpath,$ORIGIN/.."])
        os.replace(tmp, exe)
    return FDRIVE


def build_oracle(verbose: bool = False) -> None:
    """Build oracle/liboracle32.so, liboracle64.so and (if possible) oracle/_ref."""
    out = _run(["make", "-C", str(ROOT / "oracle"), "-j4"])
    if verbose:
        print(out)


def main(argv=None):
    build_library(verbose=True)
    build_fortran_host(verbose=True)
    build_oracle(verbose=True)
    print("built", LIB)


if __name__ == "__main__":
    main(sys.argv[1:])
