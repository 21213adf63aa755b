"""bench.py's contract on the CPU: a GPU count other than one is refused
with a message, and the CPU-baseline leg (the oracle on the benchmark's
model) runs and reports its sample."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_cpu_baseline_leg():
    sys.path.insert(0, ROOT)
    import bench
    from mulskips_amd import models
    out = bench.cpu_baseline(models.sic_c2_tables(), 120, 0.2, cores=2)
    assert out["kind"] == "port" and out["cores"] == 1 and out["unit"] == "events/s"
    assert out["value"] > 0 and "events in" in out["sample"]
    agg = out["aggregate"]
    assert agg["cores"] == 2 and agg["value"] > 0
