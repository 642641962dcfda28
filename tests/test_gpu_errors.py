"""GPU tests of the C ABI's argument checks: each call below must fail with
the documented status and a message, before any kernel runs, and leave the
solver usable (a valid call afterwards succeeds).

  - a time step beyond the stability limit: SM_ERR_CFL at creation, as the
    reference aborts (a_iso_3d.c:404-417, e_iso.c:476-479);
  - source and receiver cells outside the grid (3D: padded cells, 2D:
    interior cells), receiver components that cannot be recorded and values
    outside enum sourceType: SM_ERR_ARG;
  - traces asked for beyond the steps receivers were set up for: SM_ERR_ARG.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SM_ERR_ARG, SM_ERR_CFL = 1, 3
VZ, P, BULK = 2, 8, 6


@pytest.fixture(scope="module")
def ctx():
    from seismod_amd.solver import Context
    c = Context(0)
    yield c
    c.close()


def _a3d(ctx, dt=1e-3, n=12, np_=4):
    from seismod_amd.solver import A3dConfig, Acoustic3D
    n2 = n + 2 * np_
    vp = np.full((n2, n2, n2), 3000.0)
    rho = np.full((n2, n2, n2), 2000.0)
    cfg = A3dConfig(nx=n, ny=n, nz=n, np=np_, dx=10.0, dy=10.0, dz=10.0, dt=dt, f0=15.0)
    return Acoustic3D(ctx, cfg, vp, rho), n2 ** 3


def _e2d(ctx, dt=5e-4, n=16, npml=4):
    from seismod_amd.solver import E2dConfig, Elastic2D
    vp = np.full((n, n), 3000.0)
    cfg = E2dConfig(nx=n, nz=n, npml=npml, h=5.0, dt=dt)
    return Elastic2D(ctx, cfg, vp, vp / 1.8, np.full((n, n), 2000.0)), n * n   # interior cells


def _status(fn):
    from seismod_amd._lib import SeismodError
    with pytest.raises(SeismodError) as e:
        fn()
    assert str(e.value)                         # carries the library's message
    return e.value.status


def test_a3d_time_step_beyond_stability_limit(ctx):
    # limit dx / (vp sqrt(3)) = 10 / (3000 * 1.732) = 1.92 ms
    assert _status(lambda: _a3d(ctx, dt=2.5e-3)) == SM_ERR_CFL
    s, _ = _a3d(ctx, dt=1.5e-3)
    s.close()


def test_e2d_time_step_beyond_stability_limit(ctx):
    # limit h / (sqrt(2) vp) = 5 / (1.414 * 3000) = 1.18 ms
    assert _status(lambda: _e2d(ctx, dt=1.5e-3)) == SM_ERR_CFL
    s, _ = _e2d(ctx, dt=1e-3)
    s.close()


def test_a3d_cells_and_components_are_checked(ctx):
    from seismod_amd.solver import source_functions
    s, ncell = _a3d(ctx)
    L, f = source_functions([1.0], [15.0], 1e-3)
    assert _status(lambda: s.set_sources(np.array([ncell], np.uintp), L, f)) == SM_ERR_ARG
    assert _status(lambda: s.set_receivers(np.array([ncell + 5], np.uintp),
                                           np.array([VZ], np.intc), 10)) == SM_ERR_ARG
    assert _status(lambda: s.set_receivers(np.array([7], np.uintp),
                                           np.array([5], np.intc), 10)) == SM_ERR_ARG
    # valid calls still work, and traces stop at the steps set up for
    s.set_sources(np.array([ncell // 2], np.uintp), L, f)
    s.set_receivers(np.array([ncell // 2 + 3], np.uintp), np.array([P], np.intc), 10)
    s.run(10)
    assert s.traces(0, 10).shape == (10, 1)
    assert _status(lambda: s.traces(5, 6)) == SM_ERR_ARG
    s.close()


def test_e2d_cells_and_types_are_checked(ctx):
    from seismod_amd.solver import source_functions
    s, ncell = _e2d(ctx)
    L, f = source_functions([1.0], [25.0], 5e-4)
    bulk = np.array([BULK], np.intc)
    assert _status(lambda: s.set_sources(bulk, np.array([ncell], np.uintp), L, f)) == SM_ERR_ARG
    assert _status(lambda: s.set_sources(np.array([99], np.intc), np.array([5], np.uintp), L,
                                         f)) == SM_ERR_ARG
    assert _status(lambda: s.set_receivers(np.array([ncell], np.uintp),
                                           np.array([VZ], np.intc), 10)) == SM_ERR_ARG
    s.set_sources(bulk, np.array([ncell // 2], np.uintp), L, f)
    s.set_receivers(np.array([ncell // 2 + 1], np.uintp), np.array([VZ], np.intc), 8)
    s.run(8)
    assert s.traces(0, 8).shape == (8, 1)
    assert _status(lambda: s.traces(0, 9)) == SM_ERR_ARG
    s.close()
