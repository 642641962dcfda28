"""Batched 2D shots: the 64-shot e_iso workload run as one batch and as
plain per-shot launches, on one GPU.

    python -m seismod_amd.batchbench [--count 64] [--seed 4] [--steps-scale 1.0]

Workload (seismod_amd.synth.e2d_batch_shots): grids nx ~ U{300..1200} by
nz ~ U{300..800}, 20-cell PML, 3-6 layers (Vp 2000-5000 m/s, Vs = Vp/1.7-1.9,
rho 1800-2700), Ricker 10-40 Hz, 1000-3000 steps, dt at half the CFL limit,
FP32.  Set-up (model upload, autotuning of rows per chunk: per shot for
the plain launches, for the whole batch for the batched run, then of the
shot-group budget that keeps each group's fields in the last-level cache)
is outside the timed region; each timed run starts from zero state.  Prints one JSON line:
aggregate cell-updates per second batched (hipGraph replay) and one shot
after another with plain launches, and the batch's kernel time per step.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parent.parent))


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=64)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--steps-scale", type=float, default=1.0,
                    help="multiply every shot's step count (shorter runs for profiling)")
    ap.add_argument("--graph-steps", type=int, default=32)
    ap.add_argument("--no-sequential", action="store_true")
    ap.add_argument("--group-mb", default="auto",
                    help="shot-group budget in MiB for the batched run (0: all shots in every "
                         "launch), a comma list (each timed, the best reported), or 'auto' "
                         "(autotune.tune_e2d_groups picks it during set-up)")
    args = ap.parse_args(argv)

    from seismod_amd import autotune, synth
    from seismod_amd.solver import Context, E2dBatch, E2dConfig, Elastic2D, source_functions

    shots = synth.e2d_batch_shots(count=args.count, seed=args.seed)
    steps = [max(1, int(round(s.nsteps * args.steps_scale))) for s in shots]
    ctx = Context(0)
    sims = []
    for sh, n in zip(shots, steps):
        cfg = E2dConfig(nx=sh.nx, nz=sh.nz, npml=sh.npml, h=sh.h, dt=sh.dt)
        s = Elastic2D(ctx, cfg, sh.vp, sh.vs, sh.rho)
        L, f = source_functions([1.0], [sh.f0], sh.dt)
        s.set_sources(sh.src_type, sh.src_idx, L, f)
        s.set_receivers(sh.rec_idx, sh.rec_comp, n)
        autotune.tune_e2d(s, ctx, steps=10)
        s.set_launch(s.launch()[0], args.graph_steps)
        sims.append(s)
    work = sum(sh.cells * n for sh, n in zip(shots, steps))
    alone = [s.launch()[0] for s in sims]          # per-shot choice, for the plain runs
    batch_chunk, batch_variant = autotune.tune_e2d_batch(sims, ctx)

    batch = E2dBatch(ctx, graph_steps=args.graph_steps)
    batch.set_variant(batch_variant)
    groups = {}
    if args.group_mb == "auto":
        mbs = [autotune.tune_e2d_groups(batch, sims) / 2 ** 20]
    else:
        mbs = [float(x) for x in args.group_mb.split(",") if x]
    for mb in mbs:
        batch.set_group_bytes(int(mb * 2 ** 20))
        for s in sims:
            s.reset()
        batch.run(sims, [2] * len(sims))          # warm-up: graphs, code objects
        for s in sims:
            s.reset()
        batch.set_timing(True)
        t0 = time.perf_counter()
        batch.run(sims, steps)
        wall = time.perf_counter() - t0
        kms, launches = batch.timing()
        batch.set_timing(False)
        groups[mb] = (wall, kms, launches)
        print(json.dumps({"group_mb": mb, "wall_s": round(wall, 4),
                          "gcell_updates_per_s": round(work / wall / 1e9, 2)}), flush=True)
    best_mb = min(groups, key=lambda k: groups[k][0])
    wall_b, kms, launches = groups[best_mb]
    batch.close()
    out = {"shots": len(shots), "cell_updates": work,
           "padded_cells_mean": round(sum(s.cells for s in shots) / len(shots)),
           "steps_total": sum(steps),
           "batched": {"wall_s": round(wall_b, 4), "gcell_updates_per_s": round(work / wall_b / 1e9, 2),
                       "launches": launches, "kernel_ms_per_launch": round(kms / max(launches, 1), 4),
                       "graph_steps": args.graph_steps, "rows_per_chunk": batch_chunk,
                       "kernel_variant": batch_variant, "group_mb": best_mb,
                       "by_group_mb": {str(k): round(work / v[0] / 1e9, 2) for k, v in groups.items()}},
           "alg_bytes_per_cell": round(sum(synth.algorithmic_bytes_e2d(sh.nx, sh.nz, sh.npml) * sh.cells * n
                                           for sh, n in zip(shots, steps)) / work, 3)}
    if not args.no_sequential:
        for s, c in zip(sims, alone):
            s.reset()
            s.set_launch(c, 0)
        ctx.sync()
        t0 = time.perf_counter()
        for s, n in zip(sims, steps):
            s.run(n)
        ctx.sync()
        wall_s = time.perf_counter() - t0
        out["sequential_plain"] = {"wall_s": round(wall_s, 4),
                                   "gcell_updates_per_s": round(work / wall_s / 1e9, 2)}
    for s in sims:
        s.close()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
