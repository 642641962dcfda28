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


class _Req:
    def __init__(self, rid, kind, cells, nsteps):
        self.rid, self.kind, self.cells, self.nsteps = rid, kind, cells, nsteps


def _server(policy, fill_3d, fill_2d):
    # the selection needs no device: build the object without __init__
    s = serve.Server.__new__(serve.Server)
    s.policy = policy
    s.fill = {"a3d": fill_3d, "e2d": fill_2d}
    return s


def test_srpt_selects_least_remaining_work_per_kind():
    act = [dict(req=_Req(0, "a3d", 100, 50), done=0),     # left 5000
           dict(req=_Req(1, "a3d", 100, 50), done=40),    # left 1000
           dict(req=_Req(2, "a3d", 300, 10), done=0),     # left 3000
           dict(req=_Req(3, "e2d", 10, 100), done=0),     # left 1000
           dict(req=_Req(4, "e2d", 10, 20), done=0),      # left 200
           dict(req=_Req(5, "e2d", 10, 30), done=0)]      # left 300
    s = _server("srpt", fill_3d=150, fill_2d=15)
    got = [a["req"].rid for a in s._select(act)]
    # 3D: rid 1 (100 cells), then rid 2 reaches 400 >= 150; 2D: 4 then 5
    assert got == [1, 2, 4, 5]
    s = _server("srpt", fill_3d=1, fill_2d=1)
    assert [a["req"].rid for a in s._select(act)] == [1, 4]
    s = _server("share", 1, 1)
    assert [a["req"].rid for a in s._select(act)] == [0, 1, 2, 3, 4, 5]


def test_srpt_all_ranks_both_kinds_against_one_budget():
    act = [dict(req=_Req(0, "a3d", 100, 50), done=0),     # left 5000
           dict(req=_Req(1, "a3d", 100, 50), done=40),    # left 1000
           dict(req=_Req(2, "a3d", 300, 10), done=0),     # left 3000
           dict(req=_Req(3, "e2d", 10, 100), done=0),     # left 1000
           dict(req=_Req(4, "e2d", 10, 20), done=0),      # left 200
           dict(req=_Req(5, "e2d", 10, 30), done=0)]      # left 300
    # budget = the larger fill: 4 (10 cells), 5 (20), 1 (left 1000, before
    # 3 in rid order) reaches 120 >= 120; with 121, 3 follows
    s = _server("srpt-all", fill_3d=120, fill_2d=15)
    assert [a["req"].rid for a in s._select(act)] == [4, 5, 1]
    s = _server("srpt-all", fill_3d=121, fill_2d=15)
    assert [a["req"].rid for a in s._select(act)] == [4, 5, 1, 3]
    s = _server("srpt-all", fill_3d=1, fill_2d=1)
    assert [a["req"].rid for a in s._select(act)] == [4]


def test_unknown_policy_is_rejected():
    import pytest
    with pytest.raises(ValueError):
        serve.Server(policy="fifo")


def test_admission_rejects_only_what_cannot_run():
    """A request larger than the device (or one whose grid the library
    refuses: footprint 0) is rejected on its own; one that does not fit
    while others run waits."""
    GiB = 1 << 30
    assert serve.admission(4 * GiB, free=10 * GiB, reserve=2 * GiB, n_active=3) == "admit"
    assert serve.admission(9 * GiB, free=10 * GiB, reserve=2 * GiB, n_active=3) == "wait"
    assert serve.admission(9 * GiB, free=10 * GiB, reserve=2 * GiB, n_active=0) == "reject"
    assert serve.admission(0, free=10 * GiB, reserve=2 * GiB, n_active=3) == "reject"


def test_summary_counts_rejected_requests_apart():
    rs = [serve.Result(i, "e2d", 0.0, 0.0, 1.0, 10) for i in range(4)]
    rs.append(serve.Result(9, "a3d", 0.0, 0.5, 0.5, 0, None, error="too large"))
    s = serve.summarize(rs)
    assert s["requests"] == 4 and s["rejected"] == 1 and s["p99_s"] == 1.0
