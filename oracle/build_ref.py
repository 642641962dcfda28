"""Recipe for oracle/_ref: the parts of the reference that compile from
their own sources alone, built where they lie (never copied) as test
checkers.

TEST INFRASTRUCTURE ONLY.  Of seismod, only the phase-velocity routines of
pve_vti_pml's check_model (VqP1.c, VqP2.c, VqS.c with V.h: complex.h and
nothing else) build without FFTW or netCDF, so oracle/_ref holds
libvq.so with VqP1, VqP2 and VqS.  Run from the repository root (the
seismod build calls it when the reference sources are present):

    python -m oracle.build_ref [reference_dir]
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
OUT = ROOT / "oracle" / "_ref"
SOURCES = ("VqP1.c", "VqP2.c", "VqS.c")


def build(ref: Path = Path("/root/reference")) -> Path | None:
    """Compile the phase-velocity routines into oracle/_ref/libvq.so; None
    when the reference sources are absent."""
    srcs = [ref / s for s in SOURCES]
    if not all(s.exists() for s in srcs):
        return None
    OUT.mkdir(exist_ok=True)
    out = OUT / "libvq.so"
    # ROCm's clang: gcc spends tens of minutes on these generated complex
    # expressions even at -O0, clang a fraction of a second
    subprocess.run(["/opt/rocm/llvm/bin/clang", "-std=c99", "-O1", "-ffp-contract=off", "-fPIC", "-shared",
                    f"-I{ref}", *map(str, srcs), "-lm", "-o", str(out)], check=True, timeout=600)
    return out


if __name__ == "__main__":
    print(build(Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference")))
