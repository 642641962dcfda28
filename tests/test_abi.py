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


def test_e2d_fused_size_limit_without_gpu():
    """The fused FP32 2D solver masks accesses with the offset 2^31, so its
    largest buffer resource (state parities plus split fields, 96 bytes per
    padded cell) must stay below 2 GiB: bytes_needed reports 0 for a grid
    that sm_e2d_create would refuse.  FP64 and the two-kernel FP32 form have
    no such limit."""
    from seismod_amd.solver import E2dConfig
    ok = E2dConfig(nx=4729 - 40, nz=4729 - 40, npml=20, h=5.0, dt=1e-4)
    too_big = E2dConfig(nx=4730 - 40, nz=4730 - 40, npml=20, h=5.0, dt=1e-4)
    assert 96 * 4729 ** 2 < 2 ** 31 <= 96 * 4730 ** 2
    assert ok.bytes_needed() > 96 * 4729 ** 2
    assert too_big.bytes_needed() == 0
    assert E2dConfig(nx=4690, nz=4690, npml=20, h=5.0, dt=1e-4, storage="f64").bytes_needed() > 0
    assert E2dConfig(nx=4690, nz=4690, npml=20, h=5.0, dt=1e-4, naive=True).bytes_needed() > 0
