"""GPU tests of the calibration kernel (sm_probe_stream): every access
width and the non-temporal / one-tile-per-workgroup options run and report
a time; unknown widths and stream counts are rejected."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("vec", [1, 2, 4])
@pytest.mark.parametrize("flags", [0, 16, 32, 48])
def test_probe_runs(ctx, vec, flags):
    ms = ctx.probe_stream(1 << 20, 1, 1, iters=2, vec=vec | flags)
    assert ms > 0.0


def test_probe_rejects_bad_arguments(ctx):
    from seismod_amd._lib import SeismodError
    with pytest.raises(SeismodError):
        ctx.probe_stream(1 << 20, 1, 1, iters=1, vec=3)
    with pytest.raises(SeismodError):
        ctx.probe_stream(1 << 20, 7, 7, iters=1)
    with pytest.raises(SeismodError):
        ctx.probe_stream(1 << 20 | 1, 1, 1, iters=1, vec=4)
