"""Forward-modelling request server with continuous batching on one GPU.

Requests are whole shots: a 2D elastic (e_iso) or 3D acoustic (a_iso_3d)
solve with its own model, source, receivers and step count.  The server:

  - admits a request when it has arrived and the device has room for it
    (free HBM from hipMemGetInfo minus a reserve, against the solver's
    exact device footprint from sm_*_bytes_needed), in arrival order;
  - keeps every admitted shot's wavefields resident on the device until it
    finishes;
  - advances admitted shots in ticks of `tick_steps` time steps: the 3D
    shots are queued on their own streams first (asynchronous), then the
    2D shots advance together in one batched launch per time step
    (sm_e2d_batch_run), so the device works on several shots at once.
    Which shots a tick advances is the policy: "srpt-all" (default) takes
    the admitted shots with the least remaining work (cells x steps left),
    both kinds ranked together, until about one device's worth of cells is
    in flight, so short requests finish first without leaving the device
    idle and a long shot does not share the device with shorter ones;
    "srpt" does the same per kind (each kind fills its own budget);
    "share" advances every admitted shot (processor sharing);
  - between ticks, retires finished shots (traces copied back, device
    memory freed) and admits new arrivals, so requests join and leave
    while others are running.

A shot's result does not depend on what else is running: the kernels give
the same bits alone, batched, or in any launch shape.

`run_sequential` is the comparison point: one shot at a time in arrival
order, each with plain per-step launches.

CLI:  python -m seismod_amd.serve --rate 20 --count 24 [--seed 5]
          [--policy srpt|srpt-all|share]
prints one JSON line with p50/p99 latency and aggregate Gcell-updates/s for
the server and for the sequential baseline on the same arrival trace.
"""
from __future__ import annotations

import argparse
import json
import time
from dataclasses import dataclass, field

import numpy as np

from . import synth
from .solver import (A3dConfig, Acoustic3D, Context, E2dBatch, E2dConfig, Elastic2D,
                     source_functions)

VZ = 2
POLICIES = ("srpt", "srpt-all", "share")


@dataclass
class Request:
    rid: int
    kind: str                  # "e2d" or "a3d"
    arrival: float             # seconds after the start of the trace
    nsteps: int
    storage: str
    shot: object               # synth.E2dShot or synth.A3dShot
    cfg: object = None

    @property
    def cells(self) -> int:
        return self.shot.cells


@dataclass
class Result:
    rid: int
    kind: str
    arrival: float
    admitted: float
    finished: float
    cell_updates: int
    traces: np.ndarray = field(repr=False, default=None)
    error: str = ""            # non-empty: the request was rejected (no traces)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def _a3d_random_layers(n: int, seed: int, nsteps: int, f0: float) -> synth.A3dShot:
    """Random-layered 3D model (Vp 1500-5000 m/s, Gardner density), source
    at the top centre, a 16x16 Vz receiver plane; dt at 0.47 of the CFL
    limit."""
    dx = 10.0
    dt = 0.47 / (5000.0 * np.sqrt(3.0) / dx)     # Vp never exceeds 5000 m/s
    return synth.a3d_layered_shot(n=n, np_=20, dx=dx, seed=seed, vmin=1500.0, vmax=5000.0,
                                  perturb=0.0, dt=float(dt), f0=f0, nsteps=nsteps, nrec_side=16)


def make_requests(count: int, rate: float, seed: int = 5, frac_3d: float = 0.4,
                  e2d_range=(256, 2048), a3d_range=(128, 384), steps_range=(500, 4000),
                  bf16_frac: float = 0.5) -> list[Request]:
    """Poisson arrivals at `rate` per second; 2D sizes uniform in e2d_range
    per axis, 3D cube sizes uniform in a3d_range; steps uniform; 3D shots
    store their wavefields in BF16 with probability bf16_frac (FP32
    otherwise; 2D shots are FP32)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(count):
        t += float(rng.exponential(1.0 / rate)) if rate > 0 else 0.0
        nsteps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        f0 = float(rng.uniform(10.0, 30.0))
        if rng.uniform() < frac_3d:
            n = int(rng.integers(a3d_range[0], a3d_range[1] + 1))
            shot = _a3d_random_layers(n, int(rng.integers(1 << 30)), nsteps, f0)
            storage = "bf16" if rng.uniform() < bf16_frac else "f32"
            out.append(Request(rid, "a3d", t, nsteps, storage, shot))
        else:
            nx = int(rng.integers(e2d_range[0], e2d_range[1] + 1))
            nz = int(rng.integers(e2d_range[0], e2d_range[1] + 1))
            shot = synth.e2d_batch_shots(count=1, seed=int(rng.integers(1 << 30)),
                                         nx_range=(nx, nx), nz_range=(nz, nz),
                                         steps_range=(nsteps, nsteps))[0]
            out.append(Request(rid, "e2d", t, nsteps, "f32", shot))
    return out


def _config(r: Request):
    s = r.shot
    if r.kind == "a3d":
        return A3dConfig(nx=s.nx, ny=s.ny, nz=s.nz, np=s.np, dx=s.dx, dy=s.dx, dz=s.dx,
                         dt=s.dt, f0=s.f0, storage=r.storage)
    return E2dConfig(nx=s.nx, nz=s.nz, npml=s.npml, h=s.h, dt=s.dt)


def _build(ctx: Context, r: Request, graph_steps: int):
    s = r.shot
    cfg = _config(r)
    if r.kind == "a3d":
        sim = Acoustic3D(ctx, cfg, s.vp, s.rho)
        L, fct = source_functions([1.0], [s.f0], s.dt)
        sim.set_sources(s.src_idx, L, fct)
    else:
        sim = Elastic2D(ctx, cfg, s.vp, s.vs, s.rho)
        sim.set_launch(sim.launch()[0], graph_steps)
        L, fct = source_functions([1.0], [s.f0], s.dt)
        sim.set_sources(s.src_type, s.src_idx, L, fct)
    sim.set_receivers(s.rec_idx, s.rec_comp, r.nsteps)
    return sim


def _footprint(r: Request) -> int:
    """Device bytes the request's solver holds; 0 when the library would
    refuse the grid."""
    return _config(r).bytes_needed(len(r.shot.rec_idx), r.nsteps)


def admission(footprint: int, free: int, reserve: int, n_active: int) -> str:
    """"admit" when the request fits in free memory minus the reserve,
    "wait" when it does not fit now but may once running shots retire, and
    "reject" when it cannot run at all: the library refuses its grid
    (footprint 0) or it does not fit with nothing else running."""
    if footprint <= 0:
        return "reject"
    if footprint <= free - reserve:
        return "admit"
    return "wait" if n_active else "reject"


class Server:
    """Continuous-batching server on one device (see the module docstring)."""

    def __init__(self, device: int = 0, tick_steps: int = 64, reserve_bytes: int = 2 << 30,
                 max_active: int = 256, graph_steps: int = 32, policy: str = "srpt-all",
                 fill_3d: int = 16 << 20, fill_2d: int = 8 << 20):
        if policy not in POLICIES:
            raise ValueError(f"unknown policy {policy!r}")
        self.policy = policy
        self.fill = {"a3d": fill_3d, "e2d": fill_2d}   # cells in flight per tick and kind
        self.device = device
        self.tick = tick_steps
        self.reserve = reserve_bytes
        self.max_active = max_active
        self.graph_steps = graph_steps
        self.ctx = Context(device)          # 2D batch stream, admission, memory queries
        self.batch = E2dBatch(self.ctx, graph_steps=graph_steps)
        self._streams: list[Context] = []   # idle 3D streams, reused

    def _stream(self) -> Context:
        return self._streams.pop() if self._streams else Context(self.device)

    def run(self, requests: list[Request], keep_traces: bool = True) -> list[Result]:
        """Serve `requests`; returns one Result per request (rejected ones
        carry `error` and no traces).  Solvers still live when an error
        escapes are closed before it propagates."""
        active = []     # dicts: req, sim, ctx, done, admitted
        try:
            return self._serve(requests, keep_traces, active)
        finally:
            for a in active:
                a["sim"].close()
                if a["req"].kind == "a3d":
                    self._streams.append(a["ctx"])
            active.clear()

    def _serve(self, requests, keep_traces, active) -> list[Result]:
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        results = []
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0

        def reject(r, why):
            results.append(Result(r.rid, r.kind, r.arrival, clock(), clock(), 0, None, error=why))

        while pending or active:
            now = clock()
            # admission in arrival order while the device has room
            while pending and pending[0].arrival <= now and len(active) < self.max_active:
                r = pending[0]
                free, _ = self.ctx.mem_info()
                need = _footprint(r)
                verdict = admission(need, free, self.reserve, len(active))
                if verdict == "wait":
                    break
                pending.pop(0)
                if verdict == "reject":
                    reject(r, f"needs {need} device bytes (0: grid refused by the library); "
                              f"{free - self.reserve} free with {len(active)} shots running")
                    continue
                c = self._stream() if r.kind == "a3d" else self.ctx
                try:
                    sim = _build(c, r, self.graph_steps)
                except Exception as e:     # set-up refused (SM_ERR_*): this request only
                    if r.kind == "a3d":
                        self._streams.append(c)
                    reject(r, str(e))
                    continue
                active.append(dict(req=r, sim=sim, ctx=c, done=0, admitted=clock()))
            if not active:
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - clock(), 0.05)))
                continue
            # one tick: 3D shots queued on their own streams, then the 2D batch
            run = self._select(active)
            for a in active:
                a["step"] = 0
            for a in run:
                if a["req"].kind == "a3d":
                    a["step"] = min(self.tick, a["req"].nsteps - a["done"])
                    a["sim"].run(a["step"])
            e2 = [a for a in run if a["req"].kind == "e2d"]
            if e2:
                steps = [min(self.tick, a["req"].nsteps - a["done"]) for a in e2]
                self.batch.run([a["sim"] for a in e2], steps)
                for a, n in zip(e2, steps):
                    a["step"] = n
            for a in run:
                if a["req"].kind == "a3d":
                    a["ctx"].sync()
            for a in active:
                a["done"] += a["step"]
            # retire finished shots
            still = []
            for a in active:
                r = a["req"]
                if a["done"] < r.nsteps:
                    still.append(a)
                    continue
                tr = a["sim"].traces(0, r.nsteps) if keep_traces else None
                a["sim"].close()
                if r.kind == "a3d":
                    self._streams.append(a["ctx"])
                results.append(Result(r.rid, r.kind, r.arrival, a["admitted"], clock(),
                                      r.cells * r.nsteps, tr))
            active[:] = still
        return sorted(results, key=lambda x: x.rid)

    def _select(self, active: list) -> list:
        """The shots one tick advances (see the module docstring)."""
        if self.policy == "share":
            return list(active)
        left = lambda a: a["req"].cells * (a["req"].nsteps - a["done"])
        out = []
        if self.policy == "srpt-all":
            cells = 0
            budget = max(self.fill.values())
            for a in sorted(active, key=lambda a: (left(a), a["req"].rid)):
                if cells >= budget:
                    break
                out.append(a)
                cells += a["req"].cells
            return out
        for kind in ("a3d", "e2d"):
            cells = 0
            for a in sorted((a for a in active if a["req"].kind == kind),
                            key=lambda a: (left(a), a["req"].rid)):
                if cells >= self.fill[kind]:
                    break
                out.append(a)
                cells += a["req"].cells
        return out

    def close(self) -> None:
        self.batch.close()
        for c in self._streams:
            c.close()
        self._streams = []
        self.ctx.close()


def run_sequential(requests: list[Request], device: int = 0, keep_traces: bool = True):
    """One shot at a time in arrival order, plain per-step launches."""
    ctx = Context(device)
    results = []
    t0 = time.perf_counter()
    clock = lambda: time.perf_counter() - t0
    for r in sorted(requests, key=lambda r: (r.arrival, r.rid)):
        wait = r.arrival - clock()
        if wait > 0:
            time.sleep(wait)
        adm = clock()
        sim = _build(ctx, r, graph_steps=0)
        sim.run(r.nsteps)
        tr = sim.traces(0, r.nsteps) if keep_traces else None
        sim.close()
        results.append(Result(r.rid, r.kind, r.arrival, adm, clock(), r.cells * r.nsteps, tr))
    ctx.close()
    return results


def summarize(results: list[Result]) -> dict:
    """p50/p99 latency and throughput over the served requests; rejected
    requests are counted, not timed."""
    rejected = sum(1 for r in results if r.error)
    results = [r for r in results if not r.error]
    lat = np.array([r.latency for r in results])
    span = max(r.finished for r in results) - min(r.arrival for r in results)
    work = sum(r.cell_updates for r in results)
    return {"requests": len(results), "rejected": rejected, "p50_s": round(float(np.percentile(lat, 50)), 4),
            "p99_s": round(float(np.percentile(lat, 99)), 4),
            "mean_s": round(float(lat.mean()), 4), "makespan_s": round(span, 3),
            "gcell_updates_per_s": round(work / span / 1e9, 2)}


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--rate", type=float, default=20.0, help="arrivals per second")
    ap.add_argument("--count", type=int, default=24)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--tick", type=int, default=64)
    ap.add_argument("--frac-3d", type=float, default=0.4)
    ap.add_argument("--scale", type=float, default=1.0,
                    help="multiply grid edge ranges (smaller runs for quick checks)")
    ap.add_argument("--policy", choices=POLICIES, default="srpt-all",
                    help="which admitted shots a tick advances")
    ap.add_argument("--no-baseline", action="store_true")
    args = ap.parse_args(argv)
    sc = args.scale
    reqs = make_requests(args.count, args.rate, seed=args.seed, frac_3d=args.frac_3d,
                         e2d_range=(int(256 * sc), int(2048 * sc)),
                         a3d_range=(int(128 * sc), int(384 * sc)))
    srv = Server(tick_steps=args.tick, policy=args.policy)
    # warm-up: kernels loaded, streams created
    srv.run(make_requests(2, 0.0, seed=99, e2d_range=(64, 64), a3d_range=(32, 32),
                          steps_range=(20, 20)), keep_traces=False)
    res = srv.run(reqs, keep_traces=False)
    srv.close()
    line = {"rate_per_s": args.rate, "seed": args.seed, "tick_steps": args.tick, "policy": args.policy,
            "mix": {"e2d": sum(r.kind == "e2d" for r in reqs),
                    "a3d": sum(r.kind == "a3d" for r in reqs),
                    "a3d_bf16": sum(r.kind == "a3d" and r.storage == "bf16" for r in reqs)},
            "server": summarize(res)}
    if not args.no_baseline:
        line["sequential_plain"] = summarize(run_sequential(reqs, keep_traces=False))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
