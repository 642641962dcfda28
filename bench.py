#!/usr/bin/env python3
"""Benchmark: 3D acoustic CPML FDTD (a_iso_3d) on one MI355X.

One bench step is one whole shot of the workload: zero the wavefields,
run every time step (velocity, pressure, CPML, source, receivers), and copy
the receiver traces back.  The model and geometry are resident in HBM
before the timed region starts.

Default workload (a_iso_3d 256^3): 256^3 interior nodes + 20 CPML layers
(296^3 padded cells), dx = 10 m, seeded layered Vp 1500-4500 m/s with a
Gaussian-smoothed (sigma 8 cells) +-5% perturbation, Gardner density,
15 Hz Ricker Bulk source at (1280, 1280, 100) m, 32x32 Vz receivers at
z = 50 m, dt = 0.6 ms, 1500 steps, FP32.

Prints one JSON line.  value = padded cells x time steps x shots / wall
time of the timed shots (Gcell-updates/s).  After the timed region the
same shot runs once with FP64 storage and the line states the trace error
against it (error_vs_fp64; the run fails above the stated tolerance), and
the 512^3 FP32 step is timed over a few hundred steps (fp32_512).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Gcell-updates/s & %HBM roofline, 3D acoustic CPML; p50/p99 shot latency"
HBM_PEAK_GBS = 8000.0
STATE_BYTES = {"f32": 4, "bf16": 2, "f16": 2}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="a_iso_3d",
                    choices=["a_iso_3d", "e_cyl_iso", "ve_vti_sh_pml", "ve_vti_pml", "pve_iso_pml", "pve_vti_pml"],
                    help="a_iso_3d (the headline, default), e_cyl_iso (the 2.5D cylindrical "
                         "elastic solver on a borehole model), ve_vti_sh_pml or ve_vti_pml "
                         "(viscoelastic VTI SH or P-SV, pseudo-spectral), pve_iso_pml or pve_vti_pml "
                         "(poroviscoelastic Biot, isotropic or VTI, pseudo-spectral); the last five store "
                         "FP32 or FP64")
    ap.add_argument("--hipfft", action="store_true",
                    help="pseudo-spectral workloads: every derivative through hipFFT instead of the "
                         "hand-written LDS transforms")
    ap.add_argument("--steps", type=int, default=5, help="timed shots")
    ap.add_argument("--warmup", type=int, default=1, help="untimed shots")
    ap.add_argument("--n", type=int, default=256, help="interior nodes per axis")
    ap.add_argument("--nsteps", type=int, default=None, help="time steps per shot")
    ap.add_argument("--storage", default="f32", choices=sorted(STATE_BYTES) + ["f64"])
    ap.add_argument("--mat16", action="store_true",
                    help="16-bit storage only: material (rho, K*dt) as FP16 with power-of-two "
                         "scales (tile kernels)")
    ap.add_argument("--launch", default="auto",
                    help="[variant,]rows_per_block,x_chunks, or 'auto' (the autotuner picks "
                         "them during set-up, before the timed region)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-fp64-check", action="store_true",
                    help="skip the FP64 run of the same shot that the error is stated against")
    ap.add_argument("--no-512", action="store_true",
                    help="skip the extra 512^3 FP32 measurement (target size)")
    ap.add_argument("--steps-512", type=int, default=200,
                    help="time steps of the extra 512^3 measurement")
    return ap.parse_args()


# trace error of the timed storage against the FP64 run of the same shot
# (relative L2 over all traces); the bench fails above it
ERROR_TOL = {"f32": 2e-5, "f16": 4e-2, "bf16": 3e-1}
# FP32 pseudo-spectral workloads: FFT rounding over 1000 steps of a 1M-cell grid
PSTD_F32_TOL = {"ve_vti_sh_pml": 1e-5,   # measured 1.05e-6 and 1.18e-6 at 1000 steps of the 1000^2 shot
                "ve_vti_pml": 1e-4,
                "pve_iso_pml": 1e-4,
                "pve_vti_pml": 1e-4}


def cpu_baseline(shot, nsteps_sample: int) -> dict:
    """Oracle (FP64 C restatement, the reference's OpenMP shape: 3 threads for
    the velocity sections, one for pressure) on the same model; time per step
    from the difference of two runs so set-up is excluded.  Also timed with
    every sweep split over x planes on all host threads (same values)."""
    from oracle import pyoracle as O
    cfg = dict(nx=shot.nx, ny=shot.ny, nz=shot.nz, np=shot.np, dx=shot.dx, dy=shot.dx,
               dz=shot.dx, dt=shot.dt, f0=shot.f0)
    L, fct = O.source_functions([1.0], [shot.f0], shot.dt)

    def rate(**threads):
        times = []
        for n in (2, 2 + nsteps_sample):
            t0 = time.perf_counter()
            O.a3d_run(cfg, shot.vp, shot.rho, shot.src_idx, L, fct, shot.rec_idx[:1],
                      shot.rec_comp[:1], n, want_fields=False, **threads)
            times.append(time.perf_counter() - t0)
        per_step = max(times[1] - times[0], 1e-9) / nsteps_sample
        return shot.cells / per_step / 1e9

    host = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    return {"value": rate(nthreads=3, pressure_threads=1), "unit": "Gcell-updates/s", "cores": 3,
            "kind": "port",
            "sample": f"{shot.shape[0]}^3 padded model, {nsteps_sample} time steps (difference "
                      f"of a {2 + nsteps_sample}-step and a 2-step run), FP64, OpenMP sections "
                      f"as a_iso_3d.c:585 (3 threads velocity, 1 pressure)",
            "all_cores": {"value": rate(split_threads=host), "cores": host,
                          "sample": "the same run with every sweep split over x planes"}}


# kernel launched by each variant of the fused 3D path
KERNELS = {"cell": "a3d_fused_step", "cell2": "a3d_fused_step", "tile": "a3d_tile_step",
           "tilesplit": "a3d_tile_step", "tile_nt": "a3d_tile_step", "tilesplit_nt": "a3d_tile_step"}


def load_traffic(name: str, variant: str, rows: int):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary of
    this workload, if it was measured on the same kernel variant and rows
    per block (None otherwise)."""
    p = ROOT / "profiles" / f"{name}.json"
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text())
    except (OSError, ValueError):
        return None
    if d.get("variant") != variant or d.get("rows_per_block") != rows:
        return None
    return d.get("hbm_bytes_per_launch")


def fp64_error(ctx, shot, traces, nsteps) -> dict:
    """Run the timed shot once more with FP64 storage and arithmetic (the
    path whose traces and fields equal the FP64 oracle's bit for bit) and
    state the timed run's trace error against it."""
    from seismod_amd.solver import A3dConfig, Acoustic3D, source_functions
    cfg = A3dConfig(nx=shot.nx, ny=shot.ny, nz=shot.nz, np=shot.np, dx=shot.dx, dy=shot.dx,
                    dz=shot.dx, dt=shot.dt, f0=shot.f0, storage="f64")
    sim = Acoustic3D(ctx, cfg, shot.vp, shot.rho)
    L, fct = source_functions([1.0], [shot.f0], shot.dt)
    sim.set_sources(shot.src_idx, L, fct)
    sim.set_receivers(shot.rec_idx, shot.rec_comp, nsteps)
    sim.run(nsteps)
    ref = sim.traces(0, nsteps)
    sim.close()
    err = float(np.linalg.norm(traces - ref) / np.linalg.norm(ref))
    return {"traces_rel_l2": err, "against": "FP64 run of the same shot on the GPU (the FP64 path "
            "equals the FP64 oracle bit for bit)", "traces": int(ref.shape[1]), "samples": nsteps}


def measure_512(ctx, storage: str, nsteps: int) -> dict:
    """Kernel time of the 3D step at 512^3 (+ CPML 20: 552^3 padded cells,
    seed 3, Vp 1500-5000 m/s, dt 0.55 ms), autotuned launch, over nsteps
    steps from a zeroed state."""
    from seismod_amd import autotune, synth
    from seismod_amd.solver import A3dConfig, Acoustic3D, source_functions
    shot = synth.a3d_layered_shot(n=512, seed=3, dt=5.5e-4, nsteps=nsteps, vmax=5000.0)
    cfg = A3dConfig(nx=shot.nx, ny=shot.ny, nz=shot.nz, np=shot.np, dx=shot.dx, dy=shot.dx,
                    dz=shot.dx, dt=shot.dt, f0=shot.f0, storage=storage)
    sim = Acoustic3D(ctx, cfg, shot.vp, shot.rho)
    L, fct = source_functions([1.0], [shot.f0], shot.dt)
    sim.set_sources(shot.src_idx, L, fct)
    autotune.tune_a3d(sim, ctx)
    sim.set_receivers(shot.rec_idx, shot.rec_comp, nsteps)
    sim.reset()
    ctx.sync()
    sim.set_timing(True)
    t0 = time.perf_counter()
    sim.run(nsteps)
    ctx.sync()
    wall = time.perf_counter() - t0
    kms, launches = sim.timing()
    tr = sim.traces(0, nsteps)
    variant, (tj, ch) = sim.variant(), sim.launch()
    sim.close()
    if not np.isfinite(tr).all():
        raise RuntimeError("512^3 traces are not finite")
    nx2, ny2, nz2 = shot.shape
    alg_b = synth.algorithmic_bytes_a3d(nx2, ny2, nz2, shot.np, state_bytes=STATE_BYTES[storage],
                                        mat_bytes=4)
    avg_ms = kms / max(launches, 1)
    achieved = alg_b * shot.cells / (avg_ms * 1e-3) / 1e9
    return {"workload": f"a_iso_3d 512^3 + CPML 20 (552^3 padded cells), {storage} storage, "
                        f"{nsteps} time steps timed", "value": round(shot.cells * nsteps / wall / 1e9, 3),
            "unit": "Gcell-updates/s", "kernel_ms_avg": round(avg_ms, 5),
            "roofline_frac": round(achieved / HBM_PEAK_GBS, 4), "achieved_GBs": round(achieved, 1),
            "algorithmic_bytes_per_cell": round(alg_b, 3),
            "launch": {"kernel_variant": variant, "rows_per_block": tj, "x_chunks": ch}}


def source_lambda(shot) -> float:
    """Lame lambda = rho (vp^2 - 2 vs^2) of the model at the source cell, as
    the e_cyl_iso program passes it to the Kurkjian wavelet."""
    c = int(shot.src_idx[0])
    vp, vs, rho = (float(a.reshape(-1)[c]) for a in (shot.vp, shot.vs, shot.rho))
    return rho * (vp * vp - 2.0 * vs * vs)


def cpu_baseline_cyl(shot, nsteps_sample: int) -> dict:
    """Oracle (FP64 C restatement of e_cyl_iso, one thread) on the same
    borehole model; time per step from the difference of two runs."""
    from oracle import pyoracle as O
    cfg = dict(nx=shot.nx, nz=shot.nz, np=shot.np, dx=shot.dr, dz=shot.dr, dt=shot.dt, n=shot.mode,
               alpha_max=np.pi * shot.f0)
    fct = O.kurkjian_fct(source_lambda(shot), shot.dr, shot.dr, shot.src_r, shot.mode, 1e-6,
                         shot.f0, shot.dt)
    L = np.array([len(fct)], np.uintp)
    times = []
    for n in (2, 2 + nsteps_sample):
        t0 = time.perf_counter()
        O.cyl_run(cfg, shot.vp, shot.vs, shot.rho, np.array([12], np.intc), shot.src_idx, L, fct,
                  shot.rec_idx[:1], shot.rec_comp[:1], n, want_fields=False)
        times.append(time.perf_counter() - t0)
    per_step = max(times[1] - times[0], 1e-9) / nsteps_sample
    return {"value": shot.cells / per_step / 1e9, "unit": "Gcell-updates/s", "cores": 1,
            "kind": "port",
            "sample": f"{shot.vp.shape[0]}x{shot.vp.shape[1]} padded borehole model, {nsteps_sample} "
                      f"time steps (difference of a {2 + nsteps_sample}-step and a 2-step run), FP64, "
                      f"one thread"}


def main_cyl(args) -> None:
    """The e_cyl_iso solver on the seeded borehole model (1000 x 4000 cells
    of 1 cm plus a 20-cell CPML, water column of 10 cells, 8 kHz Kurkjian
    monopole on the axis, 2000 steps); one bench step is one shot."""
    from seismod_amd import synth
    from seismod_amd.solver import Context, CylConfig, Cylindrical, kurkjian_fct
    if args.storage not in ("f32", "f64"):
        print("bench.py: e_cyl_iso stores FP32 or FP64", file=sys.stderr)
        sys.exit(2)
    nsteps = args.nsteps or 2000
    shot = synth.cyl_borehole_shot(nsteps=nsteps)
    ctx = Context(0)

    def make(storage):
        cfg = CylConfig(nx=shot.nx, nz=shot.nz, np=shot.np, dr=shot.dr, dz=shot.dr, dt=shot.dt,
                        f0=shot.f0, mode=shot.mode, storage=storage)
        sim = Cylindrical(ctx, cfg, shot.vp, shot.vs, shot.rho)
        fct = kurkjian_fct(source_lambda(shot), shot.dr, shot.dr, shot.src_r, shot.mode, 1e-6,
                           shot.f0, shot.dt)
        sim.set_sources(np.array([12], np.intc), shot.src_idx, np.array([len(fct)], np.uintp), fct)
        sim.set_receivers(shot.rec_idx, shot.rec_comp, nsteps)
        return sim

    sim = make(args.storage)

    def one_shot():
        sim.reset()
        sim.run(nsteps)
        return sim.traces(0, nsteps)

    for _ in range(args.warmup):
        one_shot()
    ctx.sync()
    sim.set_timing(True)
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        tr = one_shot()
        lat.append(time.perf_counter() - t0)
    ctx.sync()
    wall = time.perf_counter() - t_start
    kms, timed = sim.timing()
    sim.close()
    if not np.isfinite(tr).all() or np.abs(tr).max() == 0.0:
        print("bench.py: traces are not finite or all zero", file=sys.stderr)
        sys.exit(1)
    nx2, nz2 = shot.vp.shape
    elt = 4 if args.storage == "f32" else 8
    alg_b = synth.algorithmic_bytes_cyl(nx2, nz2, shot.np, elt=elt)
    avg_ms = kms / max(timed, 1)
    achieved = alg_b * shot.cells / (avg_ms * 1e-3) / 1e9
    line = {
        "metric": "Gcell-updates/s & %HBM roofline, 2.5D cylindrical elastic CPML (e_cyl_iso)",
        "value": round(shot.cells * nsteps * args.steps / wall / 1e9, 3),
        "unit": "Gcell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.storage == "f32" else "f64",
        "data": "synthetic (seeded borehole model: water column in a bedded formation)",
        "config": {"workload": f"e_cyl_iso {shot.nx}x{shot.nz} + CPML {shot.np}, dr = dz = "
                               f"{shot.dr} m, mode {shot.mode}, Kurkjian {shot.f0:g} Hz, {nsteps} "
                               f"time steps per shot, {len(shot.rec_idx)} receivers",
                   "padded_grid": [nx2, nz2], "time_steps_per_shot": nsteps, "dt_s": shot.dt,
                   "parallelism": "single GPU"},
        "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                       "p99": round(1e3 * float(np.percentile(lat, 99)), 3)},
        "roofline": {"bound": "hbm", "kernel": "cyl_velocity + cyl_stress (one step)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_cell": round(alg_b, 3),
                     "kernel_ms_avg": round(avg_ms, 5), "launches_timed": timed},
        "cpu_baseline": None,
    }
    if not args.no_fp64_check and args.storage == "f32":
        ref = make("f64")
        ref.run(nsteps)
        tr64 = ref.traces(0, nsteps)
        ref.close()
        e = float(np.linalg.norm(tr - tr64) / np.linalg.norm(tr64))
        line["error_vs_fp64"] = {"traces_rel_l2": e, "against": "FP64 run of the same shot on the "
                                 "GPU (the FP64 path equals the FP64 oracle bit for bit)",
                                 "traces": len(shot.rec_idx), "samples": nsteps,
                                 "tolerance": ERROR_TOL["f32"]}
        if e > ERROR_TOL["f32"]:
            print(json.dumps(line), flush=True)
            print(f"bench.py: trace error {e:.3e} against the FP64 run exceeds "
                  f"{ERROR_TOL['f32']}", file=sys.stderr)
            sys.exit(1)
    ctx.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_cyl(shot, args.cpu_steps)
    print(json.dumps(line), flush=True)


def cpu_baseline_pstd(kind, shot, nsteps_sample: int) -> dict:
    """Oracle (FP64 numpy restatement with pocketfft, one thread) on the same
    model; time per step from the difference of two runs."""
    from oracle import pyoracle as O
    cfg = dict(nx=shot.nx, nz=shot.nz, np=shot.np, dx=shot.h, dz=shot.h, dt=shot.dt, vmax=5000.0,
               alpha_max=np.pi * shot.f0, iwipe=1)
    L, fct = O.source_functions([1.0], [shot.f0], 0.5 * shot.dt)
    src = [(PSTD[kind]["src_type"], int(shot.src_idx[0]), fct / (shot.h * shot.h))]
    if kind == "ve_vti_sh_pml":
        from oracle import oracle_sh as S
        run = lambda n: S.sh_run(cfg, shot.mu, shot.rho, shot.gamma, shot.Q, shot.f, src, shot.rec_idx[:1],
                                 shot.rec_comp[:1], n)
    elif kind in ("pve_iso_pml", "pve_vti_pml"):
        from oracle import oracle_pve as B
        mat = (B.pve_material if kind == "pve_iso_pml" else B.pve_vti_material)(*shot.model, shot.dt)
        run = lambda n: B.pve_run(cfg, mat, src, shot.rec_idx[:1], shot.rec_comp[:1], n)
    else:
        from oracle import oracle_ve as V
        mat = V.ve_material(*shot.model)
        run = lambda n: V.ve_run(cfg, mat, src, shot.rec_idx[:1], shot.rec_comp[:1], n)
    times = []
    for n in (1, 1 + nsteps_sample):
        t0 = time.perf_counter()
        run(n)
        times.append(time.perf_counter() - t0)
    per_step = max(times[1] - times[0], 1e-9) / nsteps_sample
    nx2, nz2 = shot.nx + 2 * shot.np, shot.nz + 2 * shot.np
    return {"value": shot.cells / per_step / 1e9, "unit": "Gcell-updates/s", "cores": 1,
            "kind": "port",
            "sample": f"{nx2}x{nz2} padded model, {nsteps_sample} time steps "
                      f"(difference of a {1 + nsteps_sample}-step and a 1-step run), FP64 numpy, one thread"}


# the pseudo-spectral workloads: viscoelastic VTI SH (ve_vti_sh_pml) and P-SV
# (ve_vti_pml), poroviscoelastic isotropic (pve_iso_pml); one bench step = one
# shot of 1000 steps on 1000 x 1000 nodes
PSTD = {
    "ve_vti_sh_pml": dict(src_type=5, src_name="Syz", rec_name="Vy",
                          metric="Gcell-updates/s & %HBM roofline, 2D viscoelastic VTI SH, pseudo-spectral RK4 "
                                 "(ve_vti_sh_pml)"),
    "ve_vti_pml": dict(src_type=6, src_name="Bulk", rec_name="Vz",
                       metric="Gcell-updates/s & %HBM roofline, 2D viscoelastic VTI P-SV, pseudo-spectral RK4 "
                              "(ve_vti_pml)"),
    "pve_iso_pml": dict(src_type=6, src_name="Bulk", rec_name="Vz",
                        metric="Gcell-updates/s & %HBM roofline, 2D poroviscoelastic isotropic (Biot), "
                               "pseudo-spectral RK4 (pve_iso_pml)"),
    "pve_vti_pml": dict(src_type=6, src_name="Bulk", rec_name="Vz",
                        metric="Gcell-updates/s & %HBM roofline, 2D poroviscoelastic VTI (Biot), "
                               "pseudo-spectral RK4 (pve_vti_pml)"),
}


def main_pstd(args) -> None:
    """ve_vti_sh_pml or ve_vti_pml on a seeded layered VTI viscoelastic model,
    pve_iso_pml on a seeded layered porous rock (1000 x 1000 nodes of 5 m
    plus a 20-cell CPML, one source, 200 receivers, 1000 steps); one bench
    step is one shot."""
    from seismod_amd import synth
    from seismod_amd.solver import (Context, Poroviscoelastic, PoroviscoelasticVTI, PveConfig, ShConfig, VeConfig,
                                    ViscoelasticPSV, ViscoelasticSH, source_functions)
    kind = args.workload
    spec = PSTD[kind]
    if args.storage not in ("f32", "f64"):
        print(f"bench.py: {kind} stores FP32 or FP64", file=sys.stderr)
        sys.exit(2)
    nsteps = args.nsteps or 1000
    sh, pve, pvti = kind == "ve_vti_sh_pml", kind == "pve_iso_pml", kind == "pve_vti_pml"
    shot = (synth.sh_layered_shot(nsteps=nsteps) if sh else
            synth.pve_layered_shot(nsteps=nsteps) if pve else
            synth.pve_vti_layered_shot(nsteps=nsteps) if pvti else synth.ve_layered_shot(nsteps=nsteps))
    ctx = Context(0)

    def make(storage):
        common = dict(nx=shot.nx, nz=shot.nz, np=shot.np, dx=shot.h, dz=shot.h, dt=shot.dt, f0=shot.f0,
                      storage=storage, hipfft=args.hipfft)
        if sh:
            sim = ViscoelasticSH(ctx, ShConfig(**common), shot.mu, shot.rho, shot.gamma, shot.Q, shot.f)
        elif pve:
            sim = Poroviscoelastic(ctx, PveConfig(**common), *shot.model)
        elif pvti:
            sim = PoroviscoelasticVTI(ctx, PveConfig(**common), *shot.model)
        else:
            sim = ViscoelasticPSV(ctx, VeConfig(L=shot.L, **common), *shot.model)
        L, fct = source_functions([1.0], [shot.f0], 0.5 * shot.dt)
        sim.set_sources(np.array([spec["src_type"]], np.intc), shot.src_idx, L, fct / (shot.h * shot.h))
        sim.set_receivers(shot.rec_idx, shot.rec_comp, nsteps)
        return sim

    sim = make(args.storage)

    def one_shot():
        sim.reset()            # zero fields, memory variables and traces
        sim.run(nsteps)
        return sim.traces(0, nsteps)

    for _ in range(args.warmup):
        one_shot()
    ctx.sync()
    sim.set_timing(True)
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        tr = one_shot()
        lat.append(time.perf_counter() - t0)
    ctx.sync()
    wall = time.perf_counter() - t_start
    kms, timed = sim.timing()
    sim.close()
    if not np.isfinite(tr).all() or np.abs(tr).max() == 0.0:
        print("bench.py: traces are not finite or all zero", file=sys.stderr)
        sys.exit(1)
    elt = 4 if args.storage == "f32" else 8
    alg_b = (synth.algorithmic_bytes_sh(elt) if sh else synth.algorithmic_bytes_pve(elt) if pve else
             synth.algorithmic_bytes_pve_vti(elt) if pvti else synth.algorithmic_bytes_ve(shot.L, elt))
    avg_ms = kms / max(timed, 1)
    achieved = alg_b * shot.cells / (avg_ms * 1e-3) / 1e9
    nx2, nz2 = shot.nx + 2 * shot.np, shot.nz + 2 * shot.np
    line = {
        "metric": spec["metric"],
        "value": round(shot.cells * nsteps * args.steps / wall / 1e9, 3),
        "unit": "Gcell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if args.storage == "f32" else "f64",
        "data": "synthetic (seeded layered " + ("porous rock, Biot" if pve else "VTI porous rock, Biot" if pvti
                                                 else "VTI viscoelastic") + " model)",
        "config": {"workload": f"{kind} {shot.nx}x{shot.nz} + CPML {shot.np}, h = {shot.h} m, "
                               f"{spec['src_name']} {shot.f0:g} Hz, {nsteps} time steps per shot, "
                               f"{len(shot.rec_idx)} {spec['rec_name']} receivers"
                               + ("" if sh or pve or pvti else f", L = {shot.L}"),
                   "padded_grid": [nx2, nz2], "time_steps_per_shot": nsteps, "dt_s": shot.dt,
                   "parallelism": "single GPU"},
        "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                       "p99": round(1e3 * float(np.percentile(lat, 99)), 3)},
        "roofline": {"bound": "hbm", "kernel": "one step (4 RK stages of FFT derivatives ("
                                               + ("hipFFT" if args.hipfft else "LDS transforms")
                                               + ") and a fused pointwise kernel)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_cell": alg_b, "kernel_ms_avg": round(avg_ms, 5),
                     "launches_timed": timed},
        "cpu_baseline": None,
    }
    if not args.no_fp64_check and args.storage == "f32":
        ref = make("f64")
        ref.run(nsteps)
        tr64 = ref.traces(0, nsteps)
        ref.close()
        e = float(np.linalg.norm(tr - tr64) / np.linalg.norm(tr64))
        tol = PSTD_F32_TOL[kind]
        line["error_vs_fp64"] = {"traces_rel_l2": e, "against": "FP64 run of the same shot on the GPU "
                                 "(the FP64 path equals the FP64 oracle to FFT rounding, 1e-15)",
                                 "traces": len(shot.rec_idx), "samples": nsteps, "tolerance": tol}
        if e > tol:
            print(json.dumps(line), flush=True)
            print(f"bench.py: trace error {e:.3e} against the FP64 run exceeds {tol}", file=sys.stderr)
            sys.exit(1)
    ctx.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_pstd(kind, shot, max(1, args.cpu_steps // 5))
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(shots would shard across GPUs with no exchange)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "e_cyl_iso":
        main_cyl(args)
        return
    if args.workload in PSTD:
        main_pstd(args)
        return
    if args.storage == "f64":
        print("bench.py: the a_iso_3d bench stores f32, bf16 or f16", file=sys.stderr)
        sys.exit(2)
    from seismod_amd import synth
    from seismod_amd.solver import A3dConfig, Acoustic3D, Context, source_functions

    nsteps = args.nsteps or (1500 if args.n <= 256 else 2000)
    dt = 6.0e-4 if args.n <= 256 else 5.5e-4
    shot = synth.a3d_layered_shot(n=args.n, seed=2 if args.n <= 256 else 3, dt=dt, nsteps=nsteps,
                                  vmax=4500.0 if args.n <= 256 else 5000.0)
    ctx = Context(0)
    cfg = A3dConfig(nx=shot.nx, ny=shot.ny, nz=shot.nz, np=shot.np, dx=shot.dx, dy=shot.dx,
                    dz=shot.dx, dt=shot.dt, f0=shot.f0, storage=args.storage, mat16=args.mat16)
    sim = Acoustic3D(ctx, cfg, shot.vp, shot.rho)
    lengths, fct = source_functions([1.0], [shot.f0], shot.dt)
    sim.set_sources(shot.src_idx, lengths, fct)
    if args.launch == "auto":
        from seismod_amd import autotune
        autotune.tune_a3d(sim, ctx)       # times candidates, then resets the solver
    elif args.launch:
        parts = args.launch.split(",")
        if len(parts) == 3:
            sim.set_variant(parts[0])
            parts = parts[1:]
        sim.set_launch(int(parts[0]), int(parts[1]))

    sim.set_receivers(shot.rec_idx, shot.rec_comp, nsteps)

    def one_shot():
        sim.reset()            # zero wavefields, memory variables and traces
        sim.run(nsteps)
        return sim.traces(0, nsteps)

    for _ in range(args.warmup):
        one_shot()
    ctx.sync()
    sim.set_timing(True)
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        tr = one_shot()
        lat.append(time.perf_counter() - t0)
    ctx.sync()
    wall = time.perf_counter() - t_start
    kms, launches = sim.timing()
    sim.set_timing(False)
    if not np.isfinite(tr).all() or np.abs(tr).max() == 0.0:
        print("bench.py: traces are not finite or all zero", file=sys.stderr)
        sys.exit(1)

    cells = shot.cells
    value = cells * nsteps * args.steps / wall / 1e9
    nx2, ny2, nz2 = shot.shape
    sb = STATE_BYTES[args.storage]
    alg_b = synth.algorithmic_bytes_a3d(nx2, ny2, nz2, shot.np, state_bytes=sb,
                                        mat_bytes=2 if args.mat16 else 4)
    avg_ms = kms / max(launches, 1)
    achieved = alg_b * cells / (avg_ms * 1e-3) / 1e9
    kname = f"a3d_fused_{args.storage}{'_mat16' if args.mat16 else ''}_{args.n}"
    tj, ch = sim.launch()
    variant = sim.variant()
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.storage == "f32" else
                 f"f32 compute, {args.storage} storage" + (", fp16 material" if args.mat16 else ""),
        "data": "synthetic (seeded layered model, Gardner density)",
        "config": {
            "workload": f"a_iso_3d {args.n}^3 + CPML {shot.np}, {nsteps} time steps per shot, "
                        f"{len(shot.rec_idx)} Vz receivers, {args.storage} storage"
                        + (", FP16 material" if args.mat16 else ""),
            "padded_grid": [nx2, ny2, nz2],
            "time_steps_per_shot": nsteps,
            "dt_s": shot.dt,
            "launch": {"kernel_variant": variant, "rows_per_block": tj, "x_chunks": ch,
                       "chosen_by": "autotuner" if args.launch == "auto" else "command line"},
            "parallelism": "single GPU",
        },
        "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                       "p99": round(1e3 * float(np.percentile(lat, 99)), 3)},
        "roofline": {
            "bound": "hbm",
            "kernel": KERNELS.get(variant, variant),
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": load_traffic(kname, variant, tj),
            "algorithmic_bytes_per_cell": round(alg_b, 3),
            "kernel_ms_avg": round(avg_ms, 5),
            "launches_timed": launches,
        },
        "cpu_baseline": None,
    }
    sim.close()
    if not args.no_fp64_check and args.storage != "f64":
        line["error_vs_fp64"] = fp64_error(ctx, shot, tr, nsteps)
        line["error_vs_fp64"]["tolerance"] = ERROR_TOL[args.storage]
        if line["error_vs_fp64"]["traces_rel_l2"] > ERROR_TOL[args.storage]:
            print(json.dumps(line), flush=True)
            print(f"bench.py: trace error {line['error_vs_fp64']['traces_rel_l2']:.3e} against the "
                  f"FP64 run exceeds {ERROR_TOL[args.storage]}", file=sys.stderr)
            sys.exit(1)
    if not args.no_512 and args.n != 512:
        line["fp32_512"] = measure_512(ctx, "f32", args.steps_512)
    ctx.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(shot, args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
