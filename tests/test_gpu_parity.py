"""GPU parity: the HIP kernels against the FP64 oracle, through the C ABI.

FP64 storage runs the reference's operation order without FMA contraction,
so traces and final fields must be bit-identical to the oracle.  FP32
results are held to a relative-L2 tolerance against the same oracle run.
"""
import numpy as np
import pytest
from tolerances import FP32_REL_L2, note

import decks
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

FP32_TRACE_TOL = FP32_REL_L2   # relative L2, FP32 arithmetic vs FP64 oracle
FP32_FIELD_TOL = FP32_REL_L2


def rel_l2(a, b):
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def run_a3d_gpu(ctx, d, storage, naive=False):
    from seismod_amd.solver import A3dConfig, Acoustic3D
    from seismod_amd.solver import source_functions
    c = d["cfg"]
    cfg = A3dConfig(nx=c["nx"], ny=c["ny"], nz=c["nz"], np=c["np"], dx=c["dx"], dy=c["dy"],
                    dz=c["dz"], dt=c["dt"], f0=c["f0"], vmax=c["vmax"], pml_order=c["pml_order"],
                    Rc=c["Rc"], kappa_max=c["kappa_max"], alpha_max=c["alpha_max"],
                    storage=storage, naive=naive)
    s = Acoustic3D(ctx, cfg, d["vp"], d["rho"])
    lengths, fct = source_functions(d["A"], d["f"], c["dt"])
    s.set_sources(d["src_idx"], lengths, fct)
    s.set_receivers(d["rec_idx"], d["rec_comp"], d["nsteps"])
    s.run(d["nsteps"])
    tr = s.traces(0, d["nsteps"])
    fields = np.concatenate([s.field(k).reshape(-1) for k in (8, 0, 1, 2)])
    s.close()
    return tr, fields


def run_a3d_oracle(d):
    c = d["cfg"]
    lengths, fct = O.source_functions(d["A"], d["f"], c["dt"])
    return O.a3d_run(c, d["vp"], d["rho"], d["src_idx"], lengths, fct, d["rec_idx"],
                     d["rec_comp"], d["nsteps"])


A3D_CASES = [
    dict(nx=10, ny=8, nz=6, np_=3, seed=1),
    dict(nx=7, ny=9, nz=70, np_=4, seed=2, kappa_max=3.0),          # z spans two 64-wide tiles
    dict(nx=12, ny=5, nz=9, np_=2, seed=3, dy=7.5, alpha_max=0),    # dy != dx, alpha = 0
    dict(nx=6, ny=6, nz=6, np_=0, seed=4),                           # no absorbing layers
    dict(nx=3, ny=4, nz=2, np_=5, seed=5, src_at=(1, 1, 0)),         # layers thicker than grid
]


@pytest.mark.parametrize("case", A3D_CASES, ids=lambda c: "x".join(map(str, (c["nx"], c["ny"], c["nz"]))) + f"_np{c['np_']}")
def test_a3d_fp64_bit_exact(ctx, case):
    d = decks.a3d_deck(**case)
    tr_o, f_o = run_a3d_oracle(d)
    tr_g, f_g = run_a3d_gpu(ctx, d, "f64")
    assert np.abs(tr_o).max() > 0
    assert np.array_equal(tr_g, tr_o), f"max |diff| {np.abs(tr_g - tr_o).max()}"
    assert np.array_equal(f_g, f_o), f"max |diff| {np.abs(f_g - f_o).max()}"


@pytest.mark.parametrize("case", A3D_CASES[:3], ids=lambda c: f"seed{c['seed']}")
def test_a3d_fp32_naive_tolerance(ctx, case):
    d = decks.a3d_deck(**case)
    tr_o, f_o = run_a3d_oracle(d)
    tr_g, f_g = run_a3d_gpu(ctx, d, "f32", naive=True)
    note("a3d_naive_fp32_vs_oracle", case=str(case), traces=rel_l2(tr_g, tr_o), fields=rel_l2(f_g, f_o))
    assert rel_l2(tr_g, tr_o) < FP32_TRACE_TOL
    assert rel_l2(f_g, f_o) < FP32_FIELD_TOL


def run_e2d_gpu(ctx, d, storage):
    from seismod_amd.solver import E2dConfig, Elastic2D, source_functions
    c = d["cfg"]
    cfg = E2dConfig(nx=c["nx"], nz=c["nz"], npml=c["npml"], h=c["h"], dt=c["dt"], storage=storage,
                    q1_compat=c.get("q1_compat", True),
                    damping_scale=c.get("damping_scale", 1.0))
    s = Elastic2D(ctx, cfg, d["vp"], d["vs"], d["rho"])
    lengths, fct = source_functions(d["A"], d["f"], c["dt"])
    s.set_sources(d["src_type"], d["src_idx"], lengths, fct)
    s.set_receivers(d["rec_idx"], d["rec_comp"], d["nsteps"])
    s.run(d["nsteps"])
    tr = s.traces(0, d["nsteps"])
    fields = np.concatenate([s.field_padded(k).reshape(-1) for k in (0, 2, 5, 6, 7)])
    s.close()
    return tr, fields


def run_e2d_oracle(d):
    c = d["cfg"]
    lengths, fct = O.source_functions(d["A"], d["f"], c["dt"])
    tr, f, _ = O.e2d_run(c, d["vp"], d["vs"], d["rho"], d["src_type"], d["src_idx"], lengths, fct,
                         d["rec_idx"], d["rec_comp"], d["nsteps"])
    return tr, f


E2D_CASES = [
    dict(nx=16, nz=12, npml=3, seed=1),                       # nx > nz: corrected corner
    dict(nx=12, nz=20, npml=4, seed=2),                       # nx < nz: corner reads nx*nx-1
    dict(nx=70, nz=9, npml=2, seed=3, src_type=decks.SX),
    dict(nx=9, nz=9, npml=1, seed=4, src_type=decks.SXZ),
    dict(nx=5, nz=30, npml=6, seed=5, src_type=decks.SZ, q1_compat=False),
]


@pytest.mark.parametrize("case", E2D_CASES, ids=lambda c: f"{c['nx']}x{c['nz']}_np{c['npml']}")
def test_e2d_fp64_bit_exact(ctx, case):
    d = decks.e2d_deck(**case)
    tr_o, f_o = run_e2d_oracle(d)
    tr_g, f_g = run_e2d_gpu(ctx, d, "f64")
    assert np.abs(tr_o).max() > 0
    assert np.array_equal(tr_g, tr_o), f"max |diff| {np.abs(tr_g - tr_o).max()}"
    assert np.array_equal(f_g, f_o), f"max |diff| {np.abs(f_g - f_o).max()}"


@pytest.mark.parametrize("case", E2D_CASES[:3], ids=lambda c: f"seed{c['seed']}")
def test_e2d_fp32_tolerance(ctx, case):
    d = decks.e2d_deck(**case)
    tr_o, f_o = run_e2d_oracle(d)
    tr_g, f_g = run_e2d_gpu(ctx, d, "f32")
    note("e2d_fp32_vs_oracle", case=str(case), traces=rel_l2(tr_g, tr_o), fields=rel_l2(f_g, f_o))
    assert rel_l2(tr_g, tr_o) < FP32_TRACE_TOL
    assert rel_l2(f_g, f_o) < FP32_FIELD_TOL
