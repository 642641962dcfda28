"""Storage precision on the 512^3 acoustic workload: FP32, BF16 and FP16
(power-of-two per-field scales) wavefield storage against FP64.

    python -m seismod_amd.precision [--n 512] [--nsteps 2000] [--mid 1000]
                                    [--storages f32,bf16,f16,bf16-sr,f16-sr,bf16-m16]

(a mode "-sr" rounds the 16-bit stores stochastically, "-m16" also stores
the material as FP16 with power-of-two scales)

Workload (seismod_amd.synth.a3d_layered_shot, seed 3): n^3 cells of 10 m,
4-8 layers with Vp in [1500, 5000] m/s plus a Gaussian-smoothed (8 cells)
+-5% perturbation, Gardner density, 20-cell CPML, a 15 Hz Ricker Bulk source
at (n/2, n/2, 100 m), a 32 x 32 plane of Vz receivers at z = 50 m, dt =
0.55 ms, 2000 steps.

The reference run is the library's FP64 path (SEISMOD storage "f64": the
reference's operation order in double precision, bit-identical to the FP64
CPU restatement on every grid the parity tests cover).  Every other storage
mode runs the fused kernels with FP32 arithmetic; only the stored state is
rounded.  Reported per mode, against FP64:
  - relative L2 of all traces together, and per 500-step window;
  - median and worst per-trace relative L2 over traces holding at least 1%
    of the largest trace energy;
  - relative L2 of the pressure field at step `mid` (wavelet inside the grid);
  - the kernel variant and launch shape the tuner chose and the measured
    time per step.
Prints one progress line per run and one JSON line at the end.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parent.parent))

P = 8


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    den = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def run_mode(ctx, shot, mode: str, mid: int, tune: bool):
    """Run one storage mode; returns (traces, P field at `mid`, info)."""
    from seismod_amd import autotune
    from seismod_amd.solver import A3dConfig, Acoustic3D, source_functions
    storage = mode.split("-")[0]
    cfg = A3dConfig(nx=shot.nx, ny=shot.ny, nz=shot.nz, np=shot.np, dx=shot.dx, dy=shot.dx,
                    dz=shot.dx, dt=shot.dt, f0=shot.f0, vmax=5000.0, storage=storage,
                    sround=mode.endswith("-sr"), mat16=mode.endswith("-m16"))
    s = Acoustic3D(ctx, cfg, shot.vp, shot.rho)
    lengths, fct = source_functions([1.0], [shot.f0], shot.dt)
    s.set_sources(shot.src_idx, lengths, fct)
    s.set_receivers(shot.rec_idx, shot.rec_comp, shot.nsteps)
    info = {}
    if tune and storage != "f64":
        v, tj, ch = autotune.tune_a3d(s, ctx, variants=("tile", "tilesplit"))
        info.update(variant=v, rows_per_block=tj, x_chunks=ch)
    ctx.sync()
    t0 = time.perf_counter()
    s.run(mid)
    ctx.sync()
    t1 = time.perf_counter()
    field = s.field(P)
    ctx.sync()
    t2 = time.perf_counter()
    s.run(shot.nsteps - mid)
    ctx.sync()
    t3 = time.perf_counter()
    traces = s.traces(0, shot.nsteps)
    s.close()
    info["ms_per_step"] = round(1e3 * ((t1 - t0) + (t3 - t2)) / shot.nsteps, 4)
    return traces, field, info


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--nsteps", type=int, default=2000)
    ap.add_argument("--mid", type=int, default=1000)
    ap.add_argument("--dt", type=float, default=5.5e-4)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--storages", default="f32,bf16,f16,bf16-sr,f16-sr")
    ap.add_argument("--window", type=int, default=500)
    ap.add_argument("--no-tune", action="store_true")
    args = ap.parse_args(argv)
    if not 0 < args.mid < args.nsteps:
        sys.exit("--mid must lie strictly inside (0, nsteps)")

    from seismod_amd import synth
    from seismod_amd.solver import Context

    t0 = time.perf_counter()
    shot = synth.a3d_layered_shot(n=args.n, seed=args.seed, vmin=1500.0, vmax=5000.0,
                                  dt=args.dt, nsteps=args.nsteps)
    print(f"model {args.n}^3 (+{shot.np} CPML) built in {time.perf_counter() - t0:.1f} s",
          flush=True)
    ctx = Context(0)
    ref_tr, ref_field, ref_info = run_mode(ctx, shot, "f64", args.mid, False)
    print(f"f64: {ref_info['ms_per_step']} ms/step", flush=True)
    energy = np.sum(ref_tr ** 2, axis=0)
    live = energy >= 0.01 * energy.max()
    out = {"workload": f"a_iso_3d {args.n}^3 layered, seed {args.seed}, {args.nsteps} steps, "
                       f"dt {args.dt * 1e3:g} ms",
           "receivers": int(ref_tr.shape[1]), "live_traces": int(live.sum()),
           "mid_step": args.mid, "reference": "f64", "f64_ms_per_step": ref_info["ms_per_step"],
           "modes": {}}
    for mode in [m for m in args.storages.split(",") if m]:
        tr, field, info = run_mode(ctx, shot, mode, args.mid, not args.no_tune)
        per = [rel_l2(tr[:, r], ref_tr[:, r]) for r in np.flatnonzero(live)]
        win = {f"{a}-{min(a + args.window, args.nsteps)}": round(rel_l2(tr[a:a + args.window],
                                                                        ref_tr[a:a + args.window]), 6)
               for a in range(0, args.nsteps, args.window)}
        rec = {"traces_rel_l2": round(rel_l2(tr, ref_tr), 7),
               "trace_rel_l2_median": round(float(np.median(per)), 7),
               "trace_rel_l2_worst": round(float(np.max(per)), 7),
               "windows": win,
               "p_field_rel_l2_mid": round(rel_l2(field, ref_field), 7)}
        rec.update(info)
        out["modes"][mode] = rec
        print(f"{mode}: traces {rec['traces_rel_l2']:.3e}, field {rec['p_field_rel_l2_mid']:.3e}, "
              f"{info['ms_per_step']} ms/step", flush=True)
        del tr, field
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
