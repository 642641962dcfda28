"""GPU tests of the fused FP32 2D elastic step and of shot batches.

The fused kernel's arithmetic per cell does not depend on the launch shape
(rows per chunk, graph replay), on how a run is split, or on which other
shots share the launch, so all of those must give bit-identical results.
Against the FP64 oracle it is held to a relative-L2 tolerance, and against
the reference-order FP32 kernels to a tighter one (same precision, only the
folding of dt/h and the PML coefficients differ).
"""
import numpy as np
import pytest
from tolerances import FP32_REL_L2, note

import decks
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

TOL_ORACLE = FP32_REL_L2
TOL_NAIVE = 2e-5


def rel_l2(a, b):
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def make(ctx, d, naive=False, launch=None):
    from seismod_amd.solver import E2dConfig, Elastic2D, source_functions
    c = d["cfg"]
    cfg = E2dConfig(nx=c["nx"], nz=c["nz"], npml=c["npml"], h=c["h"], dt=c["dt"], storage="f32",
                    q1_compat=c.get("q1_compat", True), naive=naive)
    s = Elastic2D(ctx, cfg, d["vp"], d["vs"], d["rho"])
    if launch is not None:
        s.set_launch(*launch)
    lengths, fct = source_functions(d["A"], d["f"], c["dt"])
    s.set_sources(d["src_type"], d["src_idx"], lengths, fct)
    s.set_receivers(d["rec_idx"], d["rec_comp"], d["nsteps"])
    return s


def outputs(s):
    tr = s.traces(0, s.step)
    f = np.concatenate([s.field_padded(k).reshape(-1) for k in (0, 2, 5, 6, 7)])
    return tr, f


def oracle(d):
    c = d["cfg"]
    lengths, fct = O.source_functions(d["A"], d["f"], c["dt"])
    tr, f, _ = O.e2d_run(c, d["vp"], d["vs"], d["rho"], d["src_type"], d["src_idx"], lengths, fct,
                         d["rec_idx"], d["rec_comp"], d["nsteps"], nthreads=3)
    return tr, f


def deck(nx, nz, npml, seed, nsteps=120, **kw):
    d = decks.e2d_deck(nx=nx, nz=nz, npml=npml, seed=seed, nsteps=nsteps, nrec=kw.pop("nrec", 12),
                       **kw)
    return d


CASES = [
    dict(nx=16, nz=12, npml=3, seed=1),
    dict(nx=12, nz=20, npml=4, seed=2),                        # Q1 corner
    dict(nx=70, nz=9, npml=2, seed=3, src_type=decks.SX),
    dict(nx=9, nz=9, npml=1, seed=4, src_type=decks.SXZ),
    dict(nx=5, nz=30, npml=6, seed=5, src_type=decks.SZ, q1_compat=False),
    dict(nx=40, nz=104, npml=10, seed=6, nsteps=200),           # NZ = 124 = 2*62
    dict(nx=90, nz=45, npml=9, seed=7, nsteps=200),             # NZ = 63: one lane into a 2nd tile
    dict(nx=30, nz=150, npml=12, seed=8, nsteps=150, src_type=decks.SX),
]


def ids(c):
    return f"{c['nx']}x{c['nz']}_np{c['npml']}"


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_fused_vs_oracle(ctx, case):
    d = deck(**case)
    s = make(ctx, d)
    s.run(d["nsteps"])
    tr, f = outputs(s)
    tr_o, f_o = oracle(d)
    assert np.abs(tr_o).max() > 0
    e_tr, e_f = rel_l2(tr, tr_o), rel_l2(f, f_o)
    note("e2d_fused_vs_oracle", case=ids(case), traces=e_tr, fields=e_f)
    assert e_tr < TOL_ORACLE and e_f < TOL_ORACLE


@pytest.mark.parametrize("case", CASES[:6], ids=ids)
def test_fused_vs_naive_fp32(ctx, case):
    d = deck(**case)
    a = make(ctx, d)
    b = make(ctx, d, naive=True)
    a.run(d["nsteps"])
    b.run(d["nsteps"])
    (ta, fa), (tb, fb) = outputs(a), outputs(b)
    assert rel_l2(ta, tb) < TOL_NAIVE and rel_l2(fa, fb) < TOL_NAIVE


@pytest.mark.parametrize("launch", [(1, 0), (2, 4), (3, 1), (7, 32), (64, 8), (4096, 0)])
def test_launch_shape_invariance(ctx, launch):
    d = deck(nx=60, nz=70, npml=8, seed=11, nsteps=90)
    ref = make(ctx, d, launch=(8, 0))
    ref.run(d["nsteps"])
    s = make(ctx, d, launch=launch)
    s.run(d["nsteps"])
    (t0, f0), (t1, f1) = outputs(ref), outputs(s)
    assert np.array_equal(t0, t1) and np.array_equal(f0, f1)


def test_run_splitting_and_reset(ctx):
    d = deck(nx=50, nz=40, npml=5, seed=12, nsteps=100)
    a = make(ctx, d, launch=(4, 8))
    a.run(100)
    b = make(ctx, d, launch=(4, 8))
    for n in (1, 30, 17, 52):
        b.run(n)
    (ta, fa), (tb, fb) = outputs(a), outputs(b)
    assert np.array_equal(ta, tb) and np.array_equal(fa, fb)
    b.reset()
    assert b.step == 0 and not b.field(0).any()
    b.run(100)
    tc, fc = outputs(b)
    assert np.array_equal(ta, tc) and np.array_equal(fa, fc)


def test_batch_equals_alone(ctx):
    """Shots of different sizes, step counts and starting steps, batched,
    give the same bits as each run alone."""
    from seismod_amd.solver import E2dBatch
    rng = np.random.default_rng(21)
    specs = []
    for k in range(7):
        nx, nz = int(rng.integers(8, 90)), int(rng.integers(8, 140))
        specs.append(dict(nx=nx, nz=nz, npml=int(rng.integers(1, 12)), seed=100 + k,
                          nsteps=int(rng.integers(40, 160)),
                          src_type=[decks.BULK, decks.SX, decks.SZ, decks.SXZ][k % 4]))
    dks = [deck(**sp) for sp in specs]
    alone = []
    for d in dks:
        s = make(ctx, d, launch=(3, 0))
        s.run(d["nsteps"])
        alone.append(outputs(s))
        s.close()
    shots = [make(ctx, d, launch=(5, 0)) for d in dks]
    # two shots start ahead on their own
    shots[1].run(13)
    shots[4].run(40)
    pre = [0] * len(shots)
    pre[1], pre[4] = 13, 40
    for gs in (0, 8):
        if gs:
            # second pass: same batch again from reset, with graph replay
            for s, p in zip(shots, pre):
                s.reset()
                if p:
                    s.run(p)
        b = E2dBatch(ctx, graph_steps=gs)
        b.run(shots, [d["nsteps"] - p for d, p in zip(dks, pre)])
        for k, (s, d) in enumerate(zip(shots, dks)):
            assert s.step == d["nsteps"]
            tr, f = outputs(s)
            assert np.array_equal(tr, alone[k][0]), (gs, k)
            assert np.array_equal(f, alone[k][1]), (gs, k)
        b.close()


@pytest.mark.parametrize("group_kib", [1, 300, 900, 1 << 20])
def test_grouped_batch_equals_alone(ctx, group_kib):
    """Shots packed into groups by footprint (budgets below every shot, a
    few shots per group, all shots in one group) and run group after
    group give the same bits as each shot alone, with and without graph
    replay, staged and per-lane forms."""
    from seismod_amd.solver import E2dBatch
    rng = np.random.default_rng(31)
    specs = []
    for k in range(9):
        nx, nz = int(rng.integers(8, 120)), int(rng.integers(8, 150))
        specs.append(dict(nx=nx, nz=nz, npml=int(rng.integers(1, 10)), seed=200 + k,
                          nsteps=int(rng.integers(30, 140)),
                          src_type=[decks.BULK, decks.SX, decks.SZ, decks.SXZ][k % 4]))
    dks = [deck(**sp) for sp in specs]
    alone = []
    for d in dks:
        s = make(ctx, d, launch=(2, 0))
        s.run(d["nsteps"])
        alone.append(outputs(s))
        s.close()
    shots = [make(ctx, d, launch=(2, 0)) for d in dks]
    for gs, variant in ((0, "rows"), (8, "staged")):
        for s in shots:
            s.reset()
        b = E2dBatch(ctx, graph_steps=gs)
        b.set_variant(variant)
        b.set_group_bytes(group_kib * 1024)
        b.run(shots, [d["nsteps"] for d in dks])
        b.close()
        for k, (s, d) in enumerate(zip(shots, dks)):
            assert s.step == d["nsteps"]
            tr, f = outputs(s)
            assert np.array_equal(tr, alone[k][0]), (gs, k)
            assert np.array_equal(f, alone[k][1]), (gs, k)
    for s in shots:
        s.close()


def test_batch_argument_errors(ctx):
    from seismod_amd.solver import E2dBatch
    from seismod_amd._lib import SeismodError
    d = deck(nx=10, nz=10, npml=2, seed=1, nsteps=10)
    a = make(ctx, d)
    n = make(ctx, d, naive=True)
    b = E2dBatch(ctx)
    with pytest.raises(SeismodError):
        b.run([a, a], [1, 1])
    with pytest.raises(SeismodError):
        b.run([n], [1])
    b.run([a], [0])
    assert a.step == 0


@pytest.mark.parametrize("case", CASES, ids=ids)
@pytest.mark.parametrize("launch", [(1, 0), (2, 0), (5, 8), (64, 0)])
def test_staged_equals_rows(ctx, case, launch):
    """The staged kernel (rows of the next step as 16-byte chunks through
    LDS) computes the same values as the per-lane form, for every chunk
    length, grid width (rows not 16-byte aligned) and PML thickness."""
    d = deck(**case)
    a = make(ctx, d, launch=launch)
    a.run(d["nsteps"])
    ra = outputs(a)
    a.close()
    b = make(ctx, d, launch=launch)
    b.set_variant("staged")
    assert b.variant() == "staged"
    b.run(d["nsteps"])
    rb = outputs(b)
    b.close()
    assert np.abs(ra[0]).max() > 0
    assert np.array_equal(ra[0], rb[0]) and np.array_equal(ra[1], rb[1])


def test_staged_batch_equals_alone(ctx):
    """A staged batch of shots of different sizes and step counts equals
    each shot run alone with the per-lane kernel."""
    from seismod_amd.solver import E2dBatch
    specs = [dict(nx=37, nz=23, npml=4, seed=21, nsteps=60), dict(nx=16, nz=70, npml=3, seed=22, nsteps=45),
             dict(nx=51, nz=130, npml=7, seed=23, nsteps=80), dict(nx=9, nz=9, npml=1, seed=24, nsteps=30)]
    ds = [deck(**c) for c in specs]
    alone = []
    for d in ds:
        s = make(ctx, d, launch=(2, 0))
        s.run(d["nsteps"])
        alone.append(outputs(s))
        s.close()
    sims = [make(ctx, d, launch=(2, 0)) for d in ds]
    b = E2dBatch(ctx, graph_steps=8)
    b.set_variant("staged")
    b.run(sims, [d["nsteps"] for d in ds])
    b.close()
    for s, (tr, f) in zip(sims, alone):
        got = outputs(s)
        assert np.array_equal(got[0], tr) and np.array_equal(got[1], f)
        s.close()
