"""GPU tests of FP16 material (SM_A3D_MAT16) in the 3D tile kernels.

With 16-bit state the material is a third of the bytes a step moves; the
flag stores rho and K*dt as FP16 with one power-of-two scale per field.
Two kinds of check:

  - exactness of the path: on a model whose rho, vp and dt make every rho
    and K*dt = rho vp^2 dt exactly representable in 11 significant bits
    (rho = 1024 x {1, 9/8, 5/4, 3/2, 2}, vp = 1024 x {1, 5/4, 3/2, 7/4},
    dt = 2^-11 s), FP16 material holds the same numbers as FP32 material,
    so every variant, launch shape and run splitting must give traces and
    fields bit-identical to the FP32-material kernel.  A wrong element
    size, offset, scale or decode in the material path shows here;
  - accuracy on a general model: against the FP64 oracle, the trace error
    of BF16 / FP16 state with FP16 material stays within 1.5x that of the
    same state with FP32 material (the material is rounded once, the state
    every step, so the state's rounding noise dominates both).
"""
import numpy as np
import pytest
from tolerances import note

import decks
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

P = 8


def exact_deck(nx=20, ny=18, nz=70, np_=4, seed=21, nsteps=90):
    d = decks.a3d_deck(nx=nx, ny=ny, nz=nz, np_=np_, seed=seed, nsteps=nsteps, dx=4.0, f0=40.0)
    rng = np.random.default_rng(seed)
    shape = d["vp"].shape
    d["vp"] = 1024.0 * rng.choice([1.0, 1.25, 1.5, 1.75], size=shape)
    d["rho"] = 1024.0 * rng.choice([1.0, 1.125, 1.25, 1.5, 2.0], size=shape)
    d["cfg"]["dt"] = 2.0 ** -11
    return d


def run(ctx, d, storage, mat16, variant="tilesplit", launch=None, split=None, sround=False):
    from seismod_amd.solver import A3dConfig, Acoustic3D, source_functions
    c = d["cfg"]
    cfg = A3dConfig(nx=c["nx"], ny=c["ny"], nz=c["nz"], np=c["np"], dx=c["dx"], dy=c["dy"],
                    dz=c["dz"], dt=c["dt"], f0=c["f0"], vmax=c["vmax"], pml_order=c["pml_order"],
                    Rc=c["Rc"], kappa_max=c["kappa_max"], alpha_max=c["alpha_max"],
                    storage=storage, mat16=mat16, sround=sround)
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
    f = np.concatenate([s.field(k).reshape(-1) for k in (P, 0, 1, 2)])
    s.close()
    return tr, f


@pytest.mark.parametrize("variant", ["tile", "tilesplit", "tile_nt", "tilesplit_nt"])
@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_mat16_equals_fp32_material_on_representable_model(ctx, storage, variant):
    d = exact_deck()
    ref = run(ctx, d, storage, False, variant, launch=(4, 3))
    assert np.abs(ref[0]).max() > 0
    for launch, split in (((4, 3), None), ((8, 2), [1, 40, 49]), ((4, 7), [45, 45])):
        got = run(ctx, d, storage, True, variant, launch=launch, split=split)
        assert np.array_equal(got[0], ref[0]), (launch, np.abs(got[0] - ref[0]).max())
        assert np.array_equal(got[1], ref[1]), (launch, np.abs(got[1] - ref[1]).max())


@pytest.fixture(scope="module")
def general():
    d = decks.a3d_deck(nx=40, ny=36, nz=44, np_=8, seed=22, nsteps=260, f0=20.0, nrec=40)
    c = d["cfg"]
    L, fct = O.source_functions(d["A"], d["f"], c["dt"])
    tr, _ = O.a3d_run(c, d["vp"], d["rho"], d["src_idx"], L, fct, d["rec_idx"], d["rec_comp"],
                      d["nsteps"], want_fields=False)
    return d, tr


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_mat16_error_vs_fp64_oracle(ctx, general, storage):
    d, tr_o = general
    rel = lambda a: float(np.linalg.norm(a - tr_o) / np.linalg.norm(tr_o))
    e32 = rel(run(ctx, d, storage, False)[0])
    e16 = rel(run(ctx, d, storage, True)[0])
    note("a3d_mat16_vs_oracle", storage=storage, fp32_material=e32, fp16_material=e16)
    # the state's rounding noise dominates both: measured 1.49e-2 against
    # 1.11e-2 (BF16) on this deck, while on the 512^3 workload BF16 with FP16
    # material measured 0.117 against 0.156 with FP32 material; a wrong
    # scale or offset in the material path gives errors of order one
    assert e16 <= 1.5 * e32 + 5e-4, (e16, e32)


def test_mat16_refusals(ctx):
    from seismod_amd._lib import SM_ERR_ARG, SeismodError
    d = exact_deck(nx=8, ny=8, nz=10, nsteps=5)
    for storage, sround in (("f32", False), ("bf16", True)):
        with pytest.raises(SeismodError) as e:
            run(ctx, d, storage, True, sround=sround)
        assert e.value.status == SM_ERR_ARG
    for variant, launch in (("cell", None), ("cell2", None), ("tile", (2, 2))):
        with pytest.raises(SeismodError) as e:
            run(ctx, d, "bf16", True, variant=variant, launch=launch)
        assert e.value.status == SM_ERR_ARG
    wide = dict(d)
    wide["vp"] = d["vp"].copy()
    wide["vp"][0, 0, 0] = 1e-7          # K*dt spanning more than 2^28
    with pytest.raises(SeismodError) as e:
        run(ctx, wide, "bf16", True)
    assert e.value.status == SM_ERR_ARG
