"""Host-side parts of the server: the request generator and the summary."""
import numpy as np

from seismod_amd import serve


def test_poisson_arrivals_and_mix_are_seeded():
    a = serve.make_requests(40, rate=10.0, seed=3, e2d_range=(16, 32), a3d_range=(8, 12),
                            steps_range=(5, 9))
    b = serve.make_requests(40, rate=10.0, seed=3, e2d_range=(16, 32), a3d_range=(8, 12),
                            steps_range=(5, 9))
    assert [(r.kind, r.arrival, r.nsteps, r.storage) for r in a] == \
           [(r.kind, r.arrival, r.nsteps, r.storage) for r in b]
    t = np.array([r.arrival for r in a])
    assert (np.diff(t) >= 0).all()
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert 0.04 < gaps.mean() < 0.25            # mean gap 1/rate = 0.1 s
    assert {r.kind for r in a} == {"e2d", "a3d"}
    for r in a:
        assert 5 <= r.nsteps <= 9
        if r.kind == "e2d":
            assert 16 <= r.shot.nx <= 32 and 16 <= r.shot.nz <= 32 and r.storage == "f32"
        else:
            assert 8 <= r.shot.nx <= 12 and r.storage in ("f32", "bf16")


def test_summary_percentiles():
    rs = [serve.Result(i, "e2d", float(i), float(i), float(i) + 1.0 + i, 10) for i in range(100)]
    s = serve.summarize(rs)
    assert s["requests"] == 100
    assert abs(s["p50_s"] - np.percentile(1.0 + np.arange(100), 50)) < 1e-9
    assert abs(s["p99_s"] - np.percentile(1.0 + np.arange(100), 99)) < 1e-9
    assert s["makespan_s"] == 99.0 + 100.0
