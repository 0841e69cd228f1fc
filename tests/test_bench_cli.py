"""bench.py's contract pieces that need no GPU: one GPU only, its flags, and
the committed traffic summary it reads for the roofline."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode != 0
    assert "one GPU" in r.stdout + r.stderr


def test_help_lists_the_workloads():
    r = run("--help")
    assert r.returncode == 0
    for w in ("icosphere", "cornell", "bdpt", "batch", "serve"):
        assert w in r.stdout


def test_traffic_summary_parses():
    sys.path.insert(0, ROOT)
    import bench
    tr = bench.measured_traffic()
    assert tr is not None
    per_launch, per_ray = tr
    assert per_launch > 1e9 and 100 < per_ray < 5000
