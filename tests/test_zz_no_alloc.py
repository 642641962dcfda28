"""GPU test of the allocation-free batch entry point, sm_e2d_batch_run_noalloc.

Called through the C ABI, it must give the same bits as sm_e2d_batch_run on
the same shots, on two calls in a row with the second one smaller (fewer,
smaller shots and fewer steps), so a stale table from the first call in the
shared device array would show.  A batch whose table exceeds the array is
refused with SM_ERR_ARG before anything runs.
"""
import ctypes as C

import numpy as np
import pytest

import decks

pytestmark = pytest.mark.gpu


def _make(ctx, d, chunk=None):
    from seismod_amd.solver import E2dConfig, Elastic2D, source_functions
    c = d["cfg"]
    s = Elastic2D(ctx, E2dConfig(nx=c["nx"], nz=c["nz"], npml=c["npml"], h=c["h"], dt=c["dt"]),
                  d["vp"], d["vs"], d["rho"])
    if chunk is not None:
        s.set_launch(chunk, 0)
    L, fct = source_functions(d["A"], d["f"], c["dt"])
    s.set_sources(d["src_type"], d["src_idx"], L, fct)
    s.set_receivers(d["rec_idx"], d["rec_comp"], d["nsteps"])
    return s


def _noalloc(batch, sims, steps):
    from seismod_amd._lib import check, lib
    arr = (C.c_void_p * len(sims))(*[s.handle.value for s in sims])
    st = np.ascontiguousarray(steps, dtype=np.uintp)
    check(lib.sm_e2d_batch_run_noalloc(batch.handle, len(sims), arr,
                                       st.ctypes.data_as(C.POINTER(C.c_size_t))),
          "sm_e2d_batch_run_noalloc")


def _outputs(s):
    return s.traces(0, s.step), np.concatenate([s.field_padded(k).reshape(-1) for k in (0, 2, 7)])


def test_noalloc_equals_batch_run_twice(ctx):
    from seismod_amd.solver import E2dBatch
    rng = np.random.default_rng(31)
    dks = []
    for k in range(5):
        dks.append(decks.e2d_deck(nx=int(rng.integers(20, 160)), nz=int(rng.integers(20, 140)),
                                  npml=int(rng.integers(2, 12)), seed=300 + k,
                                  nsteps=int(rng.integers(60, 200)), nrec=10,
                                  src_type=[decks.BULK, decks.SX, decks.SZ, decks.SXZ][k % 4]))
    calls = [(list(range(5)), None), ([1, 3], 37)]    # second call: two shots, fewer steps
    sims = [_make(ctx, d) for d in dks]
    ref = E2dBatch(ctx, graph_steps=0)
    b = E2dBatch(ctx, graph_steps=0)
    for picks, nst in calls:
        sel = [sims[i] for i in picks]
        steps = [nst or dks[i]["nsteps"] for i in picks]
        for s in sel:
            s.reset()
        ref.run(sel, steps)
        want = [_outputs(s) for s in sel]
        for s in sel:
            s.reset()
        _noalloc(b, sel, steps)
        for s, n, (tr, f) in zip(sel, steps, want):
            assert s.step == n
            got_tr, got_f = _outputs(s)
            assert np.array_equal(got_tr, tr)
            assert np.array_equal(got_f, f)
    b.close()
    ref.close()
    for s in sims:
        s.close()


def test_noalloc_refuses_a_table_beyond_the_scratch(ctx):
    """16 shots of 2048^2 (+ PML 20) at one row per chunk need a table of
    about 1.1 MB, more than the 1 MiB device array: SM_ERR_ARG, no launch."""
    from seismod_amd._lib import SM_ERR_ARG, lib
    from seismod_amd.solver import E2dBatch
    n = 2048
    d = decks.e2d_deck(nx=n, nz=n, npml=20, seed=9, nsteps=4, nrec=2, two_layer=False)
    sims = [_make(ctx, d, chunk=1) for _ in range(16)]
    b = E2dBatch(ctx, graph_steps=0)
    arr = (C.c_void_p * len(sims))(*[s.handle.value for s in sims])
    st = np.full(len(sims), 4, dtype=np.uintp)
    rc = lib.sm_e2d_batch_run_noalloc(b.handle, len(sims), arr,
                                      st.ctypes.data_as(C.POINTER(C.c_size_t)))
    assert rc == SM_ERR_ARG
    assert b"scratch" in lib.sm_last_error()
    assert all(s.step == 0 for s in sims)
    b.close()
    for s in sims:
        s.close()
