"""Bandwidth calibration with the library's plain streaming kernel.

    python -m seismod_amd.calibrate [--n N] [--iters K] [--ceiling]

Runs sm_probe_stream, whose traffic is known exactly (4 bytes x n x
(reads + writes) per launch), for a copy and for a 10-read/4-write stream
(the 3D stencil's stream count).  Two uses:
  - under `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE`, the counters of
    these launches divided by the known bytes give the factors that turn
    those counters into HBM bytes for the stencil kernels;
  - alone, it prints the bandwidth a plain streaming kernel reaches on this
    device.  With --ceiling it also runs the copy with 16-byte accesses,
    non-temporal hints and one tile per workgroup, the fastest plain copy
    found on this device: the practical ceiling next to the 8 TB/s peak.
"""
from __future__ import annotations

import argparse
import json
import sys
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parent.parent))


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=552 ** 3, help="values per array")
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--ceiling", action="store_true",
                    help="also time the fastest copy recipe (after the calibration launches)")
    args = ap.parse_args()
    from seismod_amd.solver import Context
    ctx = Context(0)
    out = []
    from seismod_amd._lib import SM_PROBE_NT, SM_PROBE_TILES
    runs = [(1, 1, 1, "persistent grid, 4 B per lane"), (10, 4, 1, "persistent grid, 4 B per lane")]
    if args.ceiling:
        runs += [(1, 1, 4 | SM_PROBE_NT | SM_PROBE_TILES, "one tile per workgroup, 16 B per lane, non-temporal"),
                 (10, 4, 4 | SM_PROBE_NT | SM_PROBE_TILES, "one tile per workgroup, 16 B per lane, non-temporal")]
    for r, w, vec, how in runs:
        ms = ctx.probe_stream(args.n, r, w, args.iters, vec=vec)
        nbytes = 4 * args.n * (r + w)
        out.append({"reads": r, "writes": w, "n": args.n, "kernel": how, "bytes_per_launch": nbytes,
                    "read_bytes": 4 * args.n * r, "write_bytes": 4 * args.n * w,
                    "ms": round(ms, 4), "GBps": round(nbytes / ms / 1e6, 1)})
        print(json.dumps(out[-1]), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
