"""CPU tests of the command-line contracts: bench.py runs on one GPU and
says so for any other count before touching a device; the seeded
model-file writer produces the header and line count the readers expect."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and not r.stdout.strip()


def test_synth_writes_a_model_file(tmp_path):
    out = tmp_path / "m.dat"
    r = subprocess.run([sys.executable, "-m", "seismod_amd.synth", "a3d", "--n", "8", "--seed", "2",
                        "--out", str(out)], cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = out.read_text().splitlines()
    assert lines[:4] == ["8 8 8", "10 10 10", "0 0 0", "isotropic_acoustic_3d"]
    assert len(lines) == 4 + 8 ** 3
    vp, rho = (float(x) for x in lines[4].split())
    assert 1500.0 * 0.9 < vp < 4500.0 * 1.1 and abs(rho - 310.0 * vp ** 0.25) < 1e-6 * rho
