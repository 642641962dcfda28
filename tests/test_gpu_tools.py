"""GPU smoke tests of the measurement programs on tiny cases: the storage
precision study (seismod_amd.precision) and the 2D batch benchmark
(seismod_amd.batchbench) run end to end and print a well-formed line."""
import json

import pytest

pytestmark = pytest.mark.gpu


def _last_json(capsys):
    out = capsys.readouterr().out.strip().splitlines()
    return json.loads(out[-1])


def test_precision_study_small(capsys):
    from seismod_amd import precision
    precision.main(["--n", "24", "--nsteps", "60", "--mid", "30", "--storages", "f32,bf16",
                    "--window", "30", "--no-tune"])
    r = _last_json(capsys)
    assert r["reference"] == "f64" and set(r["modes"]) == {"f32", "bf16"}
    f32, bf16 = r["modes"]["f32"], r["modes"]["bf16"]
    # FP32 storage stays close to FP64; BF16 is coarser but bounded
    assert f32["traces_rel_l2"] < 1e-4 and f32["p_field_rel_l2_mid"] < 1e-4
    assert f32["traces_rel_l2"] < bf16["traces_rel_l2"] < 0.5
    assert set(f32["windows"]) == {"0-30", "30-60"}


def test_batch_benchmark_small(capsys):
    from seismod_amd import batchbench
    batchbench.main(["--count", "4", "--steps-scale", "0.01", "--graph-steps", "4"])
    r = _last_json(capsys)
    assert r["shots"] == 4 and r["cell_updates"] > 0
    assert r["batched"]["gcell_updates_per_s"] > 0 and r["sequential_plain"]["gcell_updates_per_s"] > 0
    assert r["batched"]["kernel_variant"] in ("rows", "staged")
