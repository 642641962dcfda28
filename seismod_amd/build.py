"""Build recipe for libseismod_hip.so, the drop-in programs and the oracle.

Everything is built in-tree so the artefacts travel with the source tree:
  seismod_amd/lib/libseismod_hip.so   HIP kernels + C ABI (hipcc, gfx950)
  seismod_amd/bin/e_iso, a_iso_3d, e_cyl_iso, ve_vti_sh_pml, ve_vti_pml, pve_iso_pml, pve_vti_pml
                                       drop-in programs (gcc, C99)
  oracle/_build/liboracle.so           FP64 CPU restatement (test checker)

Run ``python -m seismod_amd.build`` or call :func:`build_all`.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "seismod_amd"
CSRC = PKG / "csrc"
HOST = PKG / "host"
INCLUDE = ROOT / "include"
LIBDIR = PKG / "lib"
BINDIR = PKG / "bin"
OBJDIR = PKG / "_build"
ORACLE = ROOT / "oracle"

ARCH = os.environ.get("SEISMOD_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no implicit FMA, so the FP64 kernels keep the
# reference's rounding and FP32 kernels use fma only where written.
# -packed-fp32-ops: packed FP32 VALU forms are disabled for all HIP code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}", f"-I{CSRC}",
]
C_FLAGS = ["-std=c99", "-O2", "-ffp-contract=off", "-Wall", "-Wno-unused-result"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")


def _hip_object(src: Path) -> Path:
    OBJDIR.mkdir(exist_ok=True)
    obj = OBJDIR / (src.stem + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    return obj


def build_lib() -> Path:
    """Compile every .hip file under csrc/ and link libseismod_hip.so."""
    sources = sorted(CSRC.glob("*.hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_hip_object, sources))
    LIBDIR.mkdir(exist_ok=True)
    out = LIBDIR / "libseismod_hip.so"
    tmp = out.with_suffix(".so.tmp")
    rocm_lib = Path(HIPCC).resolve().parent.parent / "lib"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), f"-L{rocm_lib}",
          "-lhipfft", f"-Wl,-rpath,{rocm_lib}", "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_host() -> list[Path]:
    """Drop-in e_iso, a_iso_3d, e_cyl_iso, ve_vti_sh_pml, ve_vti_pml, pve_iso_pml and pve_vti_pml programs linked
    against the library, and the host I/O layer alone as libsm_hostio.so (tested on CPU)."""
    BINDIR.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    _run(["gcc", *C_FLAGS, "-fopenmp", "-fPIC", "-shared", f"-I{HOST}", str(HOST / "sm_io.c"),
          "-lm", "-o", str(LIBDIR / "libsm_hostio.so")])
    programs = ("e_iso", "a_iso_3d", "e_cyl_iso", "ve_vti_sh_pml", "ve_vti_pml", "pve_iso_pml", "pve_vti_pml")
    common = sorted(p for p in HOST.glob("*.c") if p.stem not in programs)
    outs = []
    for prog in programs:
        src = HOST / f"{prog}.c"
        if not src.exists():
            continue
        out = BINDIR / prog
        _run(["gcc", *C_FLAGS, "-fopenmp", f"-I{INCLUDE}", f"-I{HOST}", str(src), *map(str, common),
              f"-L{LIBDIR}", "-lseismod_hip", "-lm",
              "-Wl,-rpath,$ORIGIN/../lib", "-o", str(out)])
        outs.append(out)
    return outs


def build_oracle() -> Path:
    """FP64 CPU restatement used as the test checker."""
    bdir = ORACLE / "_build"
    bdir.mkdir(exist_ok=True)
    out = bdir / "liboracle.so"
    srcs = sorted(ORACLE.glob("oracle_*.c"))
    _run(["gcc", *C_FLAGS, "-fopenmp", "-fPIC", "-shared", *map(str, srcs), "-lm", "-o", str(out)])
    return out


def build_ref() -> Path | None:
    """oracle/_ref: the reference's own sources that build alone (checkers
    of the tests), when the reference is present."""
    import importlib
    return importlib.import_module("oracle.build_ref").build()


def build_all() -> None:
    build_lib()
    build_host()
    build_oracle()
    build_ref()


if __name__ == "__main__":
    build_all()
    print("built", LIBDIR / "libseismod_hip.so")
