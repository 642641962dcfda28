"""GPU test of the continuous-batching server: every request's traces are
the same (bit for bit) whether it runs alone or in the mix, where 2D shots
share batched launches, 3D shots share the device on their own streams,
and requests join and leave between ticks."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("policy", ["srpt", "srpt-all", "share"])
def test_results_alone_equal_results_in_mix(policy):
    from seismod_amd import serve
    reqs = serve.make_requests(10, rate=400.0, seed=7, frac_3d=0.4, e2d_range=(40, 120),
                               a3d_range=(24, 48), steps_range=(40, 160), bf16_frac=0.5)
    kinds = {r.kind for r in reqs}
    assert kinds == {"e2d", "a3d"}
    srv = serve.Server(tick_steps=16, graph_steps=8, policy=policy)
    mix = srv.run(reqs)
    live, built = srv.batch.graphs()
    assert live <= min(built, 16)        # the graph cache stays bounded
    srv.close()
    alone = serve.run_sequential(reqs)
    assert [r.rid for r in mix] == [r.rid for r in alone] == sorted(r.rid for r in reqs)
    for a, b, q in zip(mix, alone, sorted(reqs, key=lambda r: r.rid)):
        assert a.traces.shape == (q.nsteps, len(q.shot.rec_idx))
        assert np.isfinite(a.traces).all() and np.abs(a.traces).max() > 0, q.rid
        assert np.array_equal(a.traces, b.traces), (q.rid, q.kind, q.storage)
    # requests overlapped: some request was admitted before another finished
    assert any(x.admitted < y.finished and y.admitted < x.finished
               for x in mix for y in mix if x.rid != y.rid)


def test_request_that_cannot_run_is_rejected_alone():
    """A 2D request whose grid the fused solver refuses (over 2 GiB of state)
    comes back as a rejected result; the requests around it are served."""
    from types import SimpleNamespace
    from seismod_amd import serve
    reqs = serve.make_requests(3, rate=0.0, seed=8, frac_3d=0.0, e2d_range=(40, 60),
                               steps_range=(20, 40))
    big = SimpleNamespace(nx=4800, nz=4800, npml=20, h=5.0, dt=5e-4, rec_idx=[0],
                          cells=4840 * 4840)
    reqs.append(serve.Request(99, "e2d", 0.0, 100, "f32", big))
    srv = serve.Server(tick_steps=16, graph_steps=8)
    res = srv.run(reqs)
    srv.close()
    by = {r.rid: r for r in res}
    assert by[99].error and by[99].traces is None
    assert all(not by[q.rid].error and by[q.rid].traces is not None for q in reqs[:3])
