"""bench.py's contract pieces that need no GPU: one GPU only, its flags, and
the committed counter summaries it reads for the roofline rows."""
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


def test_counter_summaries_parse():
    sys.path.insert(0, ROOT)
    import bench
    for name, wl in bench.WORKLOADS.items():
        v = bench.read_summary(wl["summary"])
        assert v is not None, wl["summary"]
        for k in ("trace_ext_bytes_per_ray", "trace_any_bytes_per_ray", "shade_bytes_per_event"):
            assert 4 < v[k] < 20000, (name, k, v[k])
        for k in ("trace_ext_valu_busy", "shade_valu_busy"):
            assert 0 < v[k] <= 100, (name, k, v[k])


def test_roofline_rows_from_a_summary():
    """The rows scale the summary's bytes per unit by the run's units per
    launch, and the HBM fraction is achieved / peak."""
    sys.path.insert(0, ROOT)
    import bench
    summary = {"source": "x.txt", "trace_ext_bytes_per_ray": 800.0, "trace_ext_valu_busy": 70.0,
               "trace_ext_valu_util": 46.0, "trace_ext_valu_insts_per_ray": 150.0, "trace_ext_td_busy": 98.0}
    row = bench.kernel_row("k", "ray", units=16e6 * 10, launches=10, ms=50.0, summary=summary,
                           key="trace_ext_bytes_per_ray", alg_bytes=2500.0, valu_key="trace_ext")
    assert row["units_per_launch"] == 16_000_000 and row["avg_launch_ms"] == 5.0
    assert row["traffic"] == 800 * 16_000_000
    assert abs(row["achieved"] - 800 * 16e6 / 5e-3 / 1e9) < 0.1
    assert abs(row["frac"] - row["achieved"] / 8000.0) < 1e-4
    assert row["algorithmic"]["bytes_per_unit"] == 2500.0 and row["valu_busy"] == 0.7
    assert row["valu_insts_per_ray"] == 150.0 and row["td_busy"] == 0.98


def test_serve_bookkeeping_with_a_stand_in_library():
    """serve_run's arrival loop, latency and plain-path bookkeeping, driven
    by a stand-in for the library (no GPU): every job is done one clock tick
    after submission."""
    import types

    import numpy as np
    sys.path.insert(0, ROOT)
    import bench

    class Job:
        def __init__(self, ctx):
            self.ctx = ctx

        def status(self):
            return {"done": True, "t_done_ms": self.ctx.t + 1.0}

        def read(self):
            return np.zeros((2, 2, 4), np.float32)

        def release(self):
            pass

        def render(self, n):
            pass

    class Ctx:
        def __init__(self, **kw):
            self.t = 0.0

        def __enter__(self):
            return self

        def __exit__(self, *a):
            pass

        def scene(self, s):
            return 0

        def job(self, *a, **k):
            return Job(self)

        def submit(self, *a):
            return Job(self)

        def clock_ms(self):
            self.t += 5.0
            return self.t

        def serve(self, n):
            return 0

        def autotune(self, *a, **k):
            return {"trace_block": 256, "mrays_per_s": 1.0}

        def config(self):
            return {"max_paths": 4, "trace_block": 256, "graph": 0, "unrelated": 1, "batch_paths": 4}

        def serve_stats(self):
            return types.SimpleNamespace(rays_ext=10, rays_shadow=5, ms_trace_ext=1.0, ms_trace_shadow=1.0,
                                         ms_shade=1.0, ms_host=0.1, groups=1, refills=1, plan_stalls=0,
                                         paths_in_flight=3)

    Y = types.SimpleNamespace(Context=Ctx, gen_icosphere_room=lambda *a: types.SimpleNamespace(tris=[0] * 10),
                              params_reference=lambda **k: None, camera=lambda *a: None)
    r = bench.serve_run(Y, np, True, 20.0, 1.0, 0, 5, 2)
    assert r["jobs"] == 20 and r["plain"]["job_at_a_time_jobs"] == 5 and r["plain"]["frame_at_a_time_jobs"] == 2
    assert r["latency_ms"]["p99"] >= r["latency_ms"]["p50"] and r["plain"]["images_identical_to_served"]
    assert r["autotuned"] and r["launch"] == {"max_paths": 4, "trace_block": 256, "graph": 0, "batch_paths": 4}
    r = bench.serve_run(Y, np, False, 20.0, 1.0, 0, 0, 2, tune=False)
    assert not r["autotuned"]
    assert r["jobs"] == 64 and r["plain"]["job_at_a_time_jobs"] == 64
