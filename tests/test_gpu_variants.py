"""GPU tests of the 3D kernel variants: one cell per lane with loads two
planes ahead ("cell2"), 64-cell tiles staged through LDS in 16-byte chunks
with 16-byte state stores ("tile"), the same with the CPML work compiled in
per tile only where a layer is touched ("tilesplit"), and those two with
non-temporal state stores ("tile_nt", "tilesplit_nt").

Each performs the same per-cell arithmetic as the one-cell-per-lane variant,
only with a different mapping of cells to lanes and memory accesses, so
traces and all four fields must be bit-identical to it for every storage
type, launch shape and run splitting.  The one-cell variant is itself held
to the FP64 oracle by test_gpu_fused.py and test_gpu_precision.py; the
oracle check is repeated here for FP32.
"""
import numpy as np
import pytest
from tolerances import FP32_REL_L2, note

import decks
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

CASES = [
    dict(nx=10, ny=8, nz=6, np_=3, seed=1),
    dict(nx=7, ny=9, nz=70, np_=4, seed=2, kappa_max=3.0),      # two z tiles, odd ny2
    dict(nx=12, ny=5, nz=8, np_=2, seed=3, dy=7.5, alpha_max=0),
    dict(nx=9, ny=11, nz=126, np_=0, seed=4),                    # no CPML, three z tiles
    dict(nx=6, ny=17, nz=64, np_=5, seed=5, nsteps=90),          # nz2 = 74
]


def ids(c):
    return f"{c['nx']}x{c['ny']}x{c['nz']}_np{c['np_']}"


def run(ctx, d, storage, variant, launch=None, sround=False, split=None, pad64=None):
    from seismod_amd.solver import A3dConfig, Acoustic3D, source_functions
    c = d["cfg"]
    cfg = A3dConfig(nx=c["nx"], ny=c["ny"], nz=c["nz"], np=c["np"], dx=c["dx"], dy=c["dy"],
                    dz=c["dz"], dt=c["dt"], f0=c["f0"], vmax=c["vmax"], pml_order=c["pml_order"],
                    Rc=c["Rc"], kappa_max=c["kappa_max"], alpha_max=c["alpha_max"],
                    storage=storage, sround=sround, pad64=bool(pad64))
    s = Acoustic3D(ctx, cfg, d["vp"], d["rho"])
    s.set_variant(variant)
    if launch is not None:
        s.set_launch(*launch)
    L, fct = source_functions(d["A"], d["f"], c["dt"])
    s.set_sources(d["src_idx"], L, fct)
    s.set_receivers(d["rec_idx"], d["rec_comp"], d["nsteps"])
    for n in (split or [d["nsteps"]]):
        s.run(n)
    tr = s.traces(0, d["nsteps"])
    f = np.concatenate([s.field(k).reshape(-1) for k in (8, 0, 1, 2)])
    s.close()
    return tr, f


VARIANTS = ["cell2", "tile", "tilesplit", "tile_nt", "tilesplit_nt"]


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("storage", ["f32", "bf16", "f16"])
@pytest.mark.parametrize("case", CASES, ids=ids)
def test_variant_equals_cell(ctx, case, storage, variant):
    d = decks.a3d_deck(**case)
    assert (d["cfg"]["nz"] + 2 * d["cfg"]["np"]) % 2 == 0
    tr_c, f_c = run(ctx, d, storage, "cell")
    tr_p, f_p = run(ctx, d, storage, variant)
    assert np.abs(tr_c).max() > 0
    assert np.array_equal(tr_p, tr_c), np.abs(tr_p - tr_c).max()
    assert np.array_equal(f_p, f_c), np.abs(f_p - f_c).max()


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("launch", [(1, 1), (2, 3), (4, 7), (8, 2), (4, 17), (8, 9)])
def test_variant_launch_shapes_and_splitting(ctx, launch, variant):
    if variant.endswith("_nt") and launch[0] < 4:
        pytest.skip("the non-temporal variants are built for 4 and 8 rows per block")
    d = decks.a3d_deck(nx=26, ny=21, nz=68, np_=4, seed=7, nsteps=70)
    ref = run(ctx, d, "f32", "cell")
    got = run(ctx, d, "f32", variant, launch=launch, split=[1, 20, 49])
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_variant_stochastic_rounding_equals_cell(ctx, storage, variant):
    # the rounding bits hash the device element index: same layout for both
    d = decks.a3d_deck(nx=14, ny=12, nz=68, np_=4, seed=8, nsteps=80)
    a = run(ctx, d, storage, "cell", sround=True)
    b = run(ctx, d, storage, variant, sround=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("variant", ["tile", "tilesplit", "tilesplit_nt"])
def test_variant_fp32_vs_oracle(ctx, variant):
    d = decks.a3d_deck(nx=11, ny=10, nz=40, np_=5, seed=9, nsteps=100)
    c = d["cfg"]
    L, fct = O.source_functions(d["A"], d["f"], c["dt"])
    tr_o, f_o = O.a3d_run(c, d["vp"], d["rho"], d["src_idx"], L, fct, d["rec_idx"],
                          d["rec_comp"], d["nsteps"])
    tr, f = run(ctx, d, "f32", variant)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    note("a3d_variant_vs_oracle", variant=variant, traces=rel(tr, tr_o), fields=rel(f, f_o))
    assert rel(tr, tr_o) < FP32_REL_L2 and rel(f, f_o) < FP32_REL_L2


@pytest.mark.parametrize("storage", ["f32", "bf16"])
@pytest.mark.parametrize("shape", [(8, 8, 7, 3), (9, 13, 1, 2), (5, 6, 131, 3)],
                         ids=["nz2_13", "nz2_5", "nz2_137"])
@pytest.mark.parametrize("variant", ["tile", "tilesplit", "tile_nt"])
def test_tile_odd_rows_equal_cell(ctx, storage, shape, variant):
    """Row pitch rounded up from an odd z extent: the chunks past the last
    cell of a row and the halo chunk of the first tile are clamped."""
    nx, ny, nz, np_ = shape
    d = decks.a3d_deck(nx=nx, ny=ny, nz=nz, np_=np_, seed=13, nsteps=40)
    a = run(ctx, d, storage, "cell")
    b = run(ctx, d, storage, variant, launch=(4 if variant.endswith("_nt") else 2, 2))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("variant", ["tilesplit", "tilesplit_nt"])
@pytest.mark.parametrize("storage", ["f32", "bf16"])
@pytest.mark.parametrize("nchunks", [1, 3, 7, 14, 21])
def test_tilesplit_ranges_equal_cell(ctx, storage, nchunks, variant):
    """x chunks that start and end inside, at and across the x absorbing
    layers (chunk boundaries at every parity), y and z tiles in and out of
    the y / z layers."""
    d = decks.a3d_deck(nx=32, ny=30, nz=140, np_=5, seed=14, nsteps=30)
    a = run(ctx, d, storage, "cell")
    b = run(ctx, d, storage, variant, launch=(4 if variant.endswith("_nt") else 2, nchunks))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_unknown_variant_ids_and_nt_row_counts_are_refused(ctx):
    from seismod_amd import _lib as L
    from seismod_amd._lib import SeismodError, check, lib
    from seismod_amd.solver import A3dConfig, Acoustic3D
    d = decks.a3d_deck(nx=8, ny=8, nz=10, np_=3, seed=11, nsteps=5)
    c = d["cfg"]
    s = Acoustic3D(ctx, A3dConfig(nx=c["nx"], ny=c["ny"], nz=c["nz"], np=c["np"], dx=c["dx"],
                                  dy=c["dy"], dz=c["dz"], dt=c["dt"], f0=c["f0"], vmax=c["vmax"]),
                   d["vp"], d["rho"])
    for v in (1, 3, 4, 9, -1):
        with pytest.raises(SeismodError) as e:
            check(lib.sm_a3d_set_variant(s.handle, v), "sm_a3d_set_variant")
        assert e.value.status == L.SM_ERR_ARG
    s.set_variant("tile_nt")
    for tj in (1, 2):
        with pytest.raises(SeismodError):
            s.set_launch(tj, 2)
    s.set_launch(8, 2)
    assert s.launch() == (8, 2)
    s.close()


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_padded_rows_equal_reference_rows(ctx, storage):
    """The same kernel on padded and unpadded device rows."""
    d = decks.a3d_deck(nx=12, ny=9, nz=40, np_=4, seed=12, nsteps=60)
    a = run(ctx, d, storage, "cell", pad64=False)
    b = run(ctx, d, storage, "cell", pad64=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
