"""The C ABI library loads on a machine without a GPU and exports every
function that include/seismod_hip.h declares; the Python binding binds them
all.  No compute call is made here."""
import re
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "seismod_hip.h"
LIB = ROOT / "seismod_amd" / "lib" / "libseismod_hip.so"


def declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("sm_a3d_create", "sm_a3d_run", "sm_e2d_create", "sm_e2d_run",
                 "sm_compute_src_fct", "sm_record_due", "sm_nsteps"):
        assert must in names


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT\s+(sm_[a-z0-9_]+)\b", out))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from seismod_amd import _lib
    missing = [n for n in declared() if n not in _lib.EXPORTED]
    assert not missing, missing
    assert _lib.lib.sm_abi_version() == 1


def test_host_only_entry_points_without_gpu():
    """Set-up helpers run on the host and need no device."""
    from seismod_amd.solver import nsteps, record_due, ricker_length, source_functions
    assert nsteps(1.0, 5e-4) == 2000
    assert ricker_length(25.0, 5e-4) == 160
    assert record_due(1, 5e-4, 0.0, 1e-3) and not record_due(0, 5e-4, 0.0, 1e-3)
    L, f = source_functions([1.0], [25.0], 5e-4)
    assert L[0] == 160 and abs(abs(f).max() - 1.0) < 1e-15
