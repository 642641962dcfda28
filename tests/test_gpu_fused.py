"""GPU tests of the fused single-pass 3D kernel (FP32 arithmetic).

Its arithmetic per cell does not depend on tiling, chunking or run
splitting, so those variations must give bit-identical results; against the
FP64 oracle it is held to a relative-L2 tolerance.
"""
import numpy as np
import pytest
from tolerances import FP32_REL_L2, note

import decks
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

FUSED_TOL = FP32_REL_L2   # relative L2 vs the FP64 oracle (traces and fields)
FUSED_VS_NAIVE_TOL = 2e-5 # reciprocal vs division in FP32, same algorithm


def rel_l2(a, b):
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def make(ctx, d, storage="f32", naive=False, launch=None):
    from seismod_amd.solver import A3dConfig, Acoustic3D, source_functions
    c = d["cfg"]
    cfg = A3dConfig(nx=c["nx"], ny=c["ny"], nz=c["nz"], np=c["np"], dx=c["dx"], dy=c["dy"],
                    dz=c["dz"], dt=c["dt"], f0=c["f0"], vmax=c["vmax"], pml_order=c["pml_order"],
                    Rc=c["Rc"], kappa_max=c["kappa_max"], alpha_max=c["alpha_max"],
                    storage=storage, naive=naive)
    s = Acoustic3D(ctx, cfg, d["vp"], d["rho"])
    if launch is not None:
        s.set_launch(*launch)
    A = d.get("A_multi", d["A"])
    f = d.get("f_multi", d["f"])
    lengths, fct = source_functions(A, f, c["dt"])
    s.set_sources(d.get("src_multi", d["src_idx"]), lengths, fct)
    s.set_receivers(d["rec_idx"], d["rec_comp"], d["nsteps"])
    return s


def outputs(s, nsteps):
    tr = s.traces(0, nsteps)
    f = np.concatenate([s.field(k).reshape(-1) for k in (8, 0, 1, 2)])
    return tr, f


def oracle(d):
    c = d["cfg"]
    A = d.get("A_multi", d["A"])
    f = d.get("f_multi", d["f"])
    lengths, fct = O.source_functions(A, f, c["dt"])
    return O.a3d_run(c, d["vp"], d["rho"], d.get("src_multi", d["src_idx"]), lengths, fct,
                     d["rec_idx"], d["rec_comp"], d["nsteps"])


CASES = [
    dict(nx=10, ny=8, nz=6, np_=3, seed=1),
    dict(nx=7, ny=9, nz=70, np_=4, seed=2, kappa_max=3.0),
    dict(nx=12, ny=5, nz=9, np_=2, seed=3, dy=7.5, alpha_max=0),
    dict(nx=6, ny=6, nz=6, np_=0, seed=4),
    dict(nx=3, ny=4, nz=2, np_=5, seed=5, src_at=(1, 1, 0)),
    dict(nx=40, ny=37, nz=150, np_=6, seed=6, nsteps=120),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['nx']}x{c['ny']}x{c['nz']}_np{c['np_']}")
def test_fused_vs_oracle(ctx, case):
    d = decks.a3d_deck(**case)
    tr_o, f_o = oracle(d)
    s = make(ctx, d)
    s.run(d["nsteps"])
    tr, f = outputs(s, d["nsteps"])
    s.close()
    note("a3d_fused_vs_oracle", case=str(case), traces=rel_l2(tr, tr_o), fields=rel_l2(f, f_o))
    assert rel_l2(tr, tr_o) < FUSED_TOL
    assert rel_l2(f, f_o) < FUSED_TOL


@pytest.mark.parametrize("case", CASES[:3] + CASES[5:], ids=lambda c: f"seed{c['seed']}")
def test_fused_vs_naive_fp32(ctx, case):
    d = decks.a3d_deck(**case)
    a = make(ctx, d)
    a.run(d["nsteps"])
    tr_a, f_a = outputs(a, d["nsteps"])
    a.close()
    b = make(ctx, d, naive=True)
    b.run(d["nsteps"])
    tr_b, f_b = outputs(b, d["nsteps"])
    b.close()
    assert rel_l2(tr_a, tr_b) < FUSED_VS_NAIVE_TOL
    assert rel_l2(f_a, f_b) < FUSED_VS_NAIVE_TOL


@pytest.mark.parametrize("launch", [(1, 1), (2, 3), (4, 7), (8, 2), (4, 26)])
def test_fused_launch_shape_invariance(ctx, launch):
    d = decks.a3d_deck(nx=40, ny=37, nz=150, np_=6, seed=6, nsteps=60)
    ref = make(ctx, d, launch=(4, 1))
    ref.run(d["nsteps"])
    tr_r, f_r = outputs(ref, d["nsteps"])
    ref.close()
    s = make(ctx, d, launch=launch)
    s.run(d["nsteps"])
    tr, f = outputs(s, d["nsteps"])
    s.close()
    assert np.array_equal(tr, tr_r)
    assert np.array_equal(f, f_r)


def test_fused_run_splitting(ctx):
    d = decks.a3d_deck(nx=20, ny=18, nz=70, np_=4, seed=7, nsteps=50)
    a = make(ctx, d)
    a.run(50)
    tr_a, f_a = outputs(a, 50)
    a.close()
    b = make(ctx, d)
    for n in (1, 7, 13, 29):
        b.run(n)
    tr_b, f_b = outputs(b, 50)
    b.close()
    assert np.array_equal(tr_a, tr_b)
    assert np.array_equal(f_a, f_b)


def test_fused_reset_repeats(ctx):
    d = decks.a3d_deck(nx=20, ny=18, nz=70, np_=4, seed=8, nsteps=40)
    s = make(ctx, d)
    s.run(40)
    tr1, f1 = outputs(s, 40)
    s.reset()
    s.set_receivers(d["rec_idx"], d["rec_comp"], 40)
    s.run(40)
    tr2, f2 = outputs(s, 40)
    s.close()
    assert np.array_equal(tr1, tr2)
    assert np.array_equal(f1, f2)


def test_fused_multiple_sources(ctx):
    """Several sources, two on one cell, sources in different x planes."""
    d = decks.a3d_deck(nx=16, ny=14, nz=20, np_=3, seed=9, nsteps=60)
    nx2, ny2, nz2 = d["vp"].shape
    cells = [(8, 7, 6), (8, 7, 6), (3, 10, 15), (12, 4, 9)]
    d["src_multi"] = np.array([(i * ny2 + j) * nz2 + k for i, j, k in cells], dtype=np.uintp)
    d["A_multi"] = np.array([1.0, 0.5, -2.0, 1.5])
    d["f_multi"] = np.array([25.0, 30.0, 20.0, 25.0])
    tr_o, f_o = oracle(d)
    s = make(ctx, d)
    s.run(d["nsteps"])
    tr, f = outputs(s, d["nsteps"])
    s.close()
    note("a3d_fused_multi_source_vs_oracle", traces=rel_l2(tr, tr_o), fields=rel_l2(f, f_o))
    assert rel_l2(tr, tr_o) < FUSED_TOL
    assert rel_l2(f, f_o) < FUSED_TOL


def test_fused_timing_counts_launches(ctx):
    d = decks.a3d_deck(nx=20, ny=18, nz=70, np_=4, seed=10, nsteps=30)
    s = make(ctx, d)
    s.set_timing(True)
    s.run(30)
    ms, n = s.timing()
    s.close()
    assert n == 30 and ms > 0.0
