"""Tolerances of the GPU parity tests, and a log of the measured errors.

FP32 arithmetic against the FP64 oracle: every FP32 kernel (fused 3D
variants, fused 2D, reference-order FP32 kernels, drop-in programs) measures
at most 3.5e-7 relative L2 on the small decks of the parity tests; the
tests hold it to about ten times that, so a CPML, PML or source term that is
off by 1e-5 of the field fails.  Longer runs at the benchmark sizes carry
their own, measured, tolerances where they are tested.
"""
import json
import os

FP32_REL_L2 = 4e-6


def note(name: str, **values) -> None:
    """Append one JSON line of measured values to $SEISMOD_REPORT (if set),
    so a test run leaves the numbers its tolerances are set from."""
    path = os.environ.get("SEISMOD_REPORT")
    if path:
        with open(path, "a") as f:
            f.write(json.dumps(dict(test=name, **{k: float(v) if hasattr(v, "__float__") else v
                                                  for k, v in values.items()})) + "\n")
