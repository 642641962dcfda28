"""Launch-shape autotuner for the fused kernels.

The fused kernels give bit-identical results for every launch shape, so the
shape is a pure performance choice.  The tuner times a few steps of each
candidate on the solver itself (HIP-event timing of the kernel launches)
and keeps the fastest; results are cached per grid shape (and storage) for
the life of the process and, optionally, in a JSON file.

Tuning runs steps and then resets the solver (fields, memory variables,
traces and step counter back to zero), so call it right after set-up,
before the solver's real run.

  3D (Acoustic3D):  kernel variant (cell, cell2, tile, tilesplit, tile_nt, tilesplit_nt)
                    x rows_per_block x x_chunks
  2D (Elastic2D):   rows_per_chunk in {1, 2, 3, 4, 6, 8, 12, 16}
  2D batches:       rows_per_chunk and kernel form for the whole batch, then
                    the shot-group budget (tune_e2d_groups)
"""
from __future__ import annotations

import json
import os
import threading
from pathlib import Path

_LOCK = threading.Lock()
_CACHE: dict[str, tuple] = {}


def _key_a3d(sim) -> str:
    c = sim.cfg
    return (f"a3d:{c.nx + 2 * c.np}x{c.ny + 2 * c.np}x{c.nz + 2 * c.np}:{c.storage}"
            + (":mat16" if getattr(c, "mat16", False) else ""))


def _key_e2d(sim) -> str:
    # shapes within 128 cells of each other share a tuning
    c = sim.cfg
    return f"e2d:{(c.nx + 2 * c.npml) // 128}x{(c.nz + 2 * c.npml) // 128}"


def _load(path):
    if path and Path(path).exists():
        try:
            for k, v in json.loads(Path(path).read_text()).items():
                _CACHE.setdefault(k, tuple(v))
        except (OSError, ValueError):
            pass


def _save(path):
    if not path:
        return
    tmp = Path(str(path) + f".{os.getpid()}.tmp")
    tmp.write_text(json.dumps({k: list(v) for k, v in _CACHE.items()}, indent=1, sort_keys=True))
    os.replace(tmp, path)


def _time(sim, ctx, steps: int) -> float:
    sim.reset()
    sim.run(2)
    ctx.sync()
    sim.set_timing(True)
    sim.run(steps)
    ms, n = sim.timing()
    sim.set_timing(False)
    return ms / max(n, 1)


def tune_a3d(sim, ctx, steps: int = 8, cache_file=None, candidates=None,
             variants=("cell", "cell2", "tile", "tilesplit", "tile_nt", "tilesplit_nt")) -> tuple:
    """Pick the kernel variant and (rows_per_block, x_chunks) for an
    Acoustic3D solver; applies them.  Returns (variant, rows_per_block,
    x_chunks)."""
    key = _key_a3d(sim)
    with _LOCK:
        _load(cache_file)
        if key in _CACHE:
            v, tj, ch = _CACHE[key]
            sim.set_variant(v)
            sim.set_launch(tj, ch)
            return _CACHE[key]
    nx2 = sim.cfg.nx + 2 * sim.cfg.np
    if candidates is None:
        candidates = [(tj, ch) for tj in (2, 4, 8) for ch in range(2, 17)
                      if ch <= max(1, nx2 // 16)]
    best = None
    for v in variants:
        try:
            sim.set_variant(v)
        except Exception:
            continue
        for tj, ch in candidates:
            try:
                sim.set_launch(tj, ch)
            except Exception:
                continue
            t = _time(sim, ctx, steps)
            if best is None or t < best[0]:
                best = (t, v, tj, ch)
    sim.reset()
    if best is None:
        return (sim.variant(),) + sim.launch()
    choice = (best[1], best[2], best[3])
    sim.set_variant(choice[0])
    sim.set_launch(choice[1], choice[2])
    with _LOCK:
        _CACHE[key] = choice
        _save(cache_file)
    return choice


def tune_e2d(sim, ctx, steps: int = 20, cache_file=None, candidates=None) -> tuple[int, int]:
    """Pick rows_per_chunk for an Elastic2D solver (graph steps kept);
    applies it.  Returns (rows_per_chunk, graph_steps)."""
    key = _key_e2d(sim)
    gs = sim.launch()[1]
    with _LOCK:
        _load(cache_file)
        if key in _CACHE:
            sim.set_launch(_CACHE[key][0], gs)
            return _CACHE[key][0], gs
    nx2 = sim.cfg.nx + 2 * sim.cfg.npml
    if candidates is None:
        candidates = [c for c in (1, 2, 3, 4, 6, 8, 12, 16) if c <= nx2]
    best = None
    for c in candidates:
        sim.set_launch(c, 0)
        t = _time(sim, ctx, steps)
        if best is None or t < best[0]:
            best = (t, c)
    sim.reset()
    sim.set_launch(best[1], gs)
    with _LOCK:
        _CACHE[key] = (best[1],)
        _save(cache_file)
    return best[1], gs


def tune_e2d_batch(sims, ctx, steps: int = 8, candidates=(1, 2, 3, 4),
                   variants=("rows", "staged")) -> int:
    """Pick one rows_per_chunk and kernel form for a batch of Elastic2D
    solvers advanced together (sm_e2d_batch_run), timing the whole batch;
    applies the chunk to every solver and resets them.  Many shots in flight favour other chunk lengths
    than one shot alone: each wave's rows are processed one after another,
    so the chunk trades waves in flight against the halo rows every chunk
    recomputes.  Returns (rows_per_chunk, variant); pass the variant to
    E2dBatch.set_variant."""
    from seismod_amd.solver import E2dBatch
    gs = [s.launch()[1] for s in sims]
    best = None
    batch = E2dBatch(ctx, graph_steps=0)
    try:
        for v in variants:
            batch.set_variant(v)
            for c in candidates:
                for s in sims:
                    s.reset()
                    s.set_launch(c, 0)
                batch.run(sims, [2] * len(sims))
                batch.set_timing(True)
                batch.run(sims, [steps] * len(sims))
                ms, n = batch.timing()
                batch.set_timing(False)
                t = ms / max(n, 1)
                if best is None or t < best[0]:
                    best = (t, c, v)
    finally:
        batch.close()
    for s, g in zip(sims, gs):
        s.reset()
        s.set_launch(best[1], g)
    return best[1], best[2]


def tune_e2d_groups(batch, sims, steps: int = 16, candidates_mb=(0, 192, 256, 320, 384)) -> int:
    """Pick the shot-group budget of an E2dBatch (sm_e2d_batch_set_group_bytes):
    each candidate runs every shot `steps` steps, group after group, timed
    on the wall clock (groups also save launches, which kernel timing would
    not show); applies the fastest and resets the solvers.  Returns bytes."""
    import time
    best = None
    for mb in candidates_mb:
        nbytes = int(mb * 2 ** 20)
        batch.set_group_bytes(nbytes)
        for s in sims:
            s.reset()
        batch.run(sims, [2] * len(sims))
        for s in sims:
            s.reset()
        t0 = time.perf_counter()
        batch.run(sims, [steps] * len(sims))
        t = time.perf_counter() - t0
        if best is None or t < best[0]:
            best = (t, nbytes)
    for s in sims:
        s.reset()
    batch.set_group_bytes(best[1])
    return best[1]


def clear_cache() -> None:
    with _LOCK:
        _CACHE.clear()
