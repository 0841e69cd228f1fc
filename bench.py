"""Benchmark: Mrays/s of the path-tracing hot path on one MI355X.

Default workload (the headline): the ~1M-triangle 1920x1080 scene (48
icospheres of 20,480 triangles in a 10x10x10 room, 2-triangle quad light),
path tracing with NEE+MIS, max depth 8, Russian roulette from bounce 3.
A step is one batch of the wavefront path tracer: 32 frames of 1920x1080
at 1 sample per pixel each (66.4M paths) run from primary rays until every
path has ended, then folded into the progressive image; 2 steps are the 64
spp of the workload.

Other single-GPU workloads (--workload):
  cornell  the 36-triangle Cornell box, 512x512, NEE+MIS, depth 8, RR from
           bounce 3; 256 frames per step, 4 steps = 1024 spp
  bdpt     bidirectional (bdpt.cl semantics) on the Cornell box, 1280x720,
           5-vertex subpaths, MIS; 8 frames per step, 4 steps = 32 spp

  batch    64 jobs on 8 icosphere scenes (20k-2M triangles), random cameras,
           128^2-1024^2, 16-256 spp, all submitted at once to the scheduler
  serve    the same job mix arriving as a Poisson stream (--rate, --duration);
           reports jobs/s, p50/p99 latency and the plain per-job paths

Rays counted: every BVH query the GPU performs (extension, NEE shadow, MIS
any-hit, BDPT connection); the CPU baseline counts with the same rule.

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BASE_SEED = 1
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_trace_ext_1m_1080p_traffic.txt")

WORKLOADS = {
    "icosphere": dict(
        # long traversals: the closest-hit and any-hit traces of an iteration overlap on two streams
        metric="Mrays/s at 1M tris 1080p", W=1920, H=1080, frames=32, steps=2, mode=0, trace_overlap=1,
        params=dict(mis=1, max_bounces=8, rr_threshold=2),
        desc="path tracing, 48 icospheres x 20480 tris + 12-tri room + 2-tri quad light ({n} tris), 1920x1080, "
             "{f} spp per step, max depth 8, RR from bounce 3, NEE+MIS"),
    "cornell": dict(
        metric="Mrays/s Cornell box 512x512", W=512, H=512, frames=256, steps=4, mode=0,
        params=dict(mis=1, max_bounces=8, rr_threshold=2),
        desc="path tracing, Cornell box ({n} tris), 512x512, {f} spp per step, max depth 8, RR from bounce 3, "
             "NEE+MIS"),
    "bdpt": dict(
        metric="Mrays/s BDPT Cornell box 720p", W=1280, H=720, frames=8, steps=4, mode=1,
        params=dict(mis=1, bdpt_bounces=5),
        desc="bidirectional path tracing (bdpt.cl), Cornell box ({n} tris), 1280x720, {f} spp per step, "
             "5-vertex eye and light subpaths, all connections, NEE+MIS"),
}


def serving_scenes(Y, np, seed=77):
    """8 icosphere rooms with triangle counts log-uniform over 20k-2M
    (ceil(N / 20480) spheres of 20,480 triangles each)."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(2e4), np.log(2e6), 8))
    return [Y.gen_icosphere_room(100 + i, int(np.ceil(n / 20480)), 5) for i, n in enumerate(counts)]


def serving_jobs(Y, np, n_scenes, n, rate, seed=78):
    """Jobs as in the batched / served configurations: a scene, a random
    camera inside the room, resolution from {128, 256, 512, 1024}^2, spp
    log-uniform over 16-256, depth 8, RR from bounce 3, NEE+MIS; Poisson
    arrivals at `rate` per second (rate 0: all at t = 0)."""
    rng = np.random.default_rng(seed)
    params = Y.params_reference(mis=1, max_bounces=8, rr_threshold=2)
    t = 0.0
    jobs = []
    for k in range(n):
        if rate > 0:
            t += rng.exponential(1.0 / rate)
        eye = rng.uniform(-4, 4, 3)
        look = rng.uniform(-4, 4, 3) - eye
        look /= np.linalg.norm(look)
        side = np.cross(look, (0.0, 1.0, 0.0))
        side /= np.linalg.norm(side)
        up = np.cross(side, look)
        res = int(rng.choice([128, 256, 512, 1024]))
        spp = int(round(np.exp(rng.uniform(np.log(16), np.log(256)))))
        jobs.append(dict(t=t, scene=int(rng.integers(n_scenes)), cam=Y.camera(side, up, look, eye, 60.0), W=res, H=res,
                         spp=spp, seed=int(rng.integers(1, 1 << 30)), params=params))
    return jobs


def serve_bench(args, Y, np):
    """Served (continuous batching through the scheduler) vs one job at a
    time on the plain per-job path, on the same jobs (so equal spp and
    estimator: the images are identical)."""
    serve = args.workload == "serve"
    scenes = serving_scenes(Y, np)
    n_jobs = args.jobs or (int(args.duration * args.rate) if serve else 64)
    jobs = serving_jobs(Y, np, len(scenes), n_jobs, args.rate if serve else 0.0)
    lat, done_t = [], []
    with Y.Context(mem_budget=64 << 30, graph=args.graph, trace_overlap=max(args.trace_overlap, 0)) as ctx:
        t0 = time.perf_counter()
        sids = [ctx.scene(s) for s in scenes]
        print(f"# {len(scenes)} scenes, {sum(len(s.tris) for s in scenes)} triangles, uploaded in "
              f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
        # warm the kernels once
        w = ctx.job(sids[0], jobs[0]["cam"], 64, 64, jobs[0]["params"])
        w.render(1)
        w.release()
        out = {}
        clock0 = ctx.clock_ms()
        nxt = 0
        live = {}
        while nxt < len(jobs) or live:
            now = ctx.clock_ms() - clock0
            while nxt < len(jobs) and jobs[nxt]["t"] * 1e3 <= now:
                j = jobs[nxt]
                live[nxt] = ctx.submit(sids[j["scene"]], j["cam"], j["W"], j["H"], j["params"], j["spp"], j["seed"])
                nxt += 1
            if live:
                ctx.serve(1)
            elif nxt < len(jobs):
                time.sleep(max(0.0, jobs[nxt]["t"] - now / 1e3))
            for k in list(live):
                st = live[k].status()
                if st["done"]:
                    lat.append(st["t_done_ms"] - clock0 - jobs[k]["t"] * 1e3)
                    done_t.append(st["t_done_ms"] - clock0)
                    if k < 3:
                        out[k] = live[k].read()
                    live[k].release()
                    del live[k]
        makespan = max(done_t) / 1e3
        sst = ctx.serve_stats()
        # plain paths on a sample of the same jobs
        n_plain = min(len(jobs), args.plain_jobs) if args.plain_jobs > 0 else len(jobs)
        t_frame = t_batch = 0.0
        same = True
        for k in range(n_plain):
            j = jobs[k]
            a = ctx.job(sids[j["scene"]], j["cam"], j["W"], j["H"], j["params"], base_seed=j["seed"])
            t1 = time.perf_counter()
            for _ in range(j["spp"]):
                a.render(1)
            t_frame += time.perf_counter() - t1
            if k in out:
                same &= bool(np.array_equal(a.read().view(np.int32), out[k].view(np.int32)))
            a.release()
            b = ctx.job(sids[j["scene"]], j["cam"], j["W"], j["H"], j["params"], base_seed=j["seed"])
            t1 = time.perf_counter()
            b.render(j["spp"])
            b.read()
            t_batch += time.perf_counter() - t1
            b.release()
    lat = np.array(lat)
    served = len(jobs) / makespan
    plain = n_plain / t_frame
    plain_b = n_plain / t_batch
    line = {
        "metric": "served jobs/s with p50/p99 latency vs per-job" if serve else "batched jobs/s vs per-job",
        "value": round(served, 3), "unit": "jobs/s", "n_gpus": 1, "steps": len(jobs), "warmup": 1,
        "ms_per_step": round(makespan * 1e3 / len(jobs), 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {
            "workload": ("Poisson arrivals at %.1f jobs/s for %.1f s" % (args.rate, args.duration) if serve else
                         "%d jobs submitted together" % len(jobs)) +
                        ", 8 icosphere scenes (20k-2M tris), random cameras, 128^2-1024^2, 16-256 spp, depth 8, "
                        "NEE+MIS, 64 GB image budget",
            "jobs": len(jobs), "scene_tris": [len(s.tris) for s in scenes],
            "total_spp_megasamples": round(sum(j["spp"] * j["W"] * j["H"] for j in jobs) / 1e6, 1),
        },
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 1), "p99": round(float(np.percentile(lat, 99)), 1),
                       "max": round(float(lat.max()), 1)},
        "makespan_s": round(makespan, 3),
        "served_mrays_per_s": round((sst.rays_ext + sst.rays_shadow) / makespan / 1e6, 1),
        "plain": {"jobs_per_s_frame_at_a_time": round(plain, 3), "jobs_per_s_job_at_a_time": round(plain_b, 3),
                  "sample_jobs": n_plain, "speedup_vs_frame_at_a_time": round(served / plain, 2),
                  "speedup_vs_job_at_a_time": round(served / plain_b, 2),
                  "images_identical_to_served": same},
    }
    print(json.dumps(line))


def make_scene(Y, name):
    if name == "icosphere":
        return Y.gen_icosphere_room(2024, 48, 5)
    return Y.gen_cornell()


def cpu_baseline(scene, nodes, wl, threads: int, budget_s: float = 12.0):
    """The CPU restatement of udpt.cl / bdpt.cl (reference BFS traversal,
    HEAP_SIZE 1500) on the same scene, camera and seeds, over a centre tile."""
    from oracle import pyoracle as O
    W, H = wl["W"], wl["H"]
    osc = O.OScene(scene.tris, scene.mats, nodes, scene.light, heap_size=1500)
    cam = O.Cam.from_buffer_copy(bytes(scene.cam))
    p = O.params(wl["mode"], **wl["params"])
    tw, th = 128, 128
    x0, y0 = W // 2 - tw // 2, H // 2 - th // 2
    rect = (x0, y0, x0 + tw, y0 + th)
    seeds = O.mt19937(BASE_SEED, 64)
    t0 = time.perf_counter()
    osc.render(cam, W, H, p, seeds[:1], rect=rect, threads=threads)
    t1 = time.perf_counter() - t0
    spp = int(max(1, min(63, round(budget_s / max(t1, 1e-3)))))
    t0 = time.perf_counter()
    _, _, rays, samples = osc.render(cam, W, H, p, seeds[1:1 + spp], rect=rect, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(rays / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{tw}x{th} centre tile of the same {W}x{H} frame, {spp} spp (frames 2..{1 + spp}), "
                      f"{samples} samples, {rays} BVH queries in {dt:.1f} s, {threads} threads"}


def copy_probe(n_bytes: int = 1 << 30, reps: int = 3) -> None:
    """A plain device-to-device copy of known size (torch's copy kernel), for
    scaling the FETCH_SIZE / WRITE_SIZE counters: under `rocprofv3 --pmc`
    each dispatch should read and write n_bytes."""
    import torch
    src = torch.ones(n_bytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n_bytes, "dispatches": reps}))


def measured_traffic():
    """HBM bytes per closest-hit trace launch from the committed PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), or None."""
    try:
        vals = {}
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if ":" in line and not line.startswith(("#", " ")):
                    k, v = line.split(":", 1)
                    vals[k.strip()] = v.strip()
        return float(vals["traffic_bytes_per_launch"]), float(vals["traffic_bytes_per_ray"])
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed steps (0: the workload's default)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS) + ["batch", "serve"], default="icosphere")
    ap.add_argument("--rate", type=float, default=20.0, help="serve: Poisson arrival rate (jobs/s)")
    ap.add_argument("--duration", type=float, default=3.0, help="serve: seconds of arrivals")
    ap.add_argument("--jobs", type=int, default=0, help="batch / serve: number of jobs (0: 64 / rate x duration)")
    ap.add_argument("--plain-jobs", type=int, default=0, help="jobs timed on the plain per-job paths (0: all)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--bvh-width", type=int, default=0)
    ap.add_argument("--trace-block", type=int, default=0)
    ap.add_argument("--trace-chunk", type=int, default=0)
    ap.add_argument("--trace-grid", type=int, default=0)
    ap.add_argument("--lds-stack", type=int, default=0)
    ap.add_argument("--shade-waves", type=int, default=0)
    ap.add_argument("--max-paths", type=int, default=0, help="paths in flight per batch (0: library default)")
    ap.add_argument("--frames", type=int, default=0, help="frames per step (0: the workload's default)")
    ap.add_argument("--batch-factor", type=int, default=0, help="work per batch over paths in flight (2: regeneration)")
    ap.add_argument("--trace-overlap", type=int, default=-1,
                    help="1: closest-hit and any-hit traces on two streams, 0: one stream, -1: the workload's choice")
    ap.add_argument("--graph", type=int, default=0, help="1: bounce loop replayed from a HIP graph")
    ap.add_argument("--trace-xcd", type=int, default=0, help="ray-queue slices per trace launch (1: no slicing)")
    ap.add_argument("--trace-refill", type=int, default=0, help="idle lanes before a trace wave takes new rays (-1: off)")
    ap.add_argument("--leaf-max", type=int, default=0, help="triangles per BVH leaf (0: library default)")
    ap.add_argument("--bvh-build", type=int, default=0, help="GPU tree: 1 own SAH build, 2 the reference's tree")
    ap.add_argument("--hoist-frac", type=float, default=0.0, help="box-area fraction above which triangles are hoisted")
    ap.add_argument("--copy-probe", action="store_true", help="run only the copy kernel used to calibrate PMC counters")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU design is described but "
              "not built", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import yune_amd as Y

    if args.workload in ("batch", "serve"):
        serve_bench(args, Y, np)
        return
    wl = WORKLOADS[args.workload]
    W, H, F = wl["W"], wl["H"], args.frames or wl["frames"]
    steps = args.steps or wl["steps"]
    t0 = time.perf_counter()
    scene = make_scene(Y, args.workload)
    nodes = scene.bvh()
    t_build = time.perf_counter() - t0
    print(f"# scene {len(scene.tris)} triangles, {len(nodes)} BVH nodes, build {t_build:.2f}s", file=sys.stderr,
          flush=True)

    params = Y.params_reference(wl["mode"], **wl["params"])
    with Y.Context(bvh_width=args.bvh_width, trace_block=args.trace_block, trace_chunk=args.trace_chunk,
                   leaf_max=args.leaf_max, hoist_frac=args.hoist_frac,
                   trace_grid=args.trace_grid, lds_stack=args.lds_stack, shade_waves=args.shade_waves,
                   max_paths=args.max_paths, batch_factor=args.batch_factor,
                   trace_overlap=args.trace_overlap if args.trace_overlap >= 0 else wl.get("trace_overlap", 0),
                   graph=args.graph,
                   trace_xcd=args.trace_xcd or wl.get("trace_xcd", 0),
                   trace_refill=args.trace_refill or wl.get("trace_refill", 0), bvh_build=args.bvh_build) as ctx:
        sid = ctx.scene(scene, nodes)
        info = ctx.scene_info(sid)
        launch = {k: v for k, v in ctx.config().items()
                  if k in ("max_paths", "lds_stack", "shade_waves", "trace_block", "trace_chunk", "leaf_max",
                           "trace_xcd", "trace_refill", "graph", "trace_overlap", "batch_factor", "bvh_build")}
        # counting pass: exact node visits / triangle tests of one step
        cj = ctx.job(sid, scene.cam, W, H, params, base_seed=BASE_SEED)
        cj.set_counting(True)
        cj.render(F)
        cst = cj.stats()
        cj.release()
        job = ctx.job(sid, scene.cam, W, H, params, base_seed=BASE_SEED)
        for _ in range(args.warmup):
            job.render(F)
        s0 = job.stats()
        t0 = time.perf_counter()
        for _ in range(steps):
            job.render(F)
        dt = time.perf_counter() - t0
        s1 = job.stats()
        img = job.read()

    rays = (s1.rays_ext - s0.rays_ext) + (s1.rays_shadow - s0.rays_shadow)
    mrays = rays / dt / 1e6
    # roofline of the dominant kernel (closest-hit traversal): algorithmic
    # bytes per extension ray from the counting pass: interior nodes visited
    # x node size + triangle tests x 48 B + the ray (32 B) and its hit (16 B)
    node_b = 64 if info.width == 2 else 128
    ext_bytes_per_ray = (cst.nodes_ext * node_b + cst.tris_ext * 48) / max(cst.rays_ext, 1) + 32 + 16
    ext_rays = s1.rays_ext - s0.rays_ext
    ext_ms = s1.ms_trace_ext - s0.ms_trace_ext
    launches = s1.trace_launches_ext - s0.trace_launches_ext
    achieved = ext_bytes_per_ray * ext_rays / (ext_ms / 1e3) / 1e9 if ext_ms > 0 else 0.0
    line = {
        "metric": wl["metric"],
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / max(steps, 1), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": wl["desc"].format(n=len(scene.tris), f=F),
            "triangles": len(scene.tris), "bvh_nodes_ref": len(nodes), "bvh_width": info.width,
            "resolution": [W, H], "spp_per_step": F, "spp_total": F * steps,
            "rays": int(rays), "rays_per_sample": round(rays / max(s1.samples - s0.samples, 1), 3),
            "ms_trace_ext": round(ext_ms, 2), "ms_trace_shadow": round(s1.ms_trace_shadow - s0.ms_trace_shadow, 2),
            "ms_shade": round(s1.ms_shade - s0.ms_shade, 2),
            "launch": launch,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_trace (closest hit)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "traffic_per_ray": None,
            "bytes_per_ray": round(ext_bytes_per_ray, 1),
            "nodes_per_ray": round(cst.nodes_ext / max(cst.rays_ext, 1), 2),
            "tri_tests_per_ray": round(cst.tris_ext / max(cst.rays_ext, 1), 2),
            "rays_per_launch": round(ext_rays / max(launches, 1)),
            "avg_launch_ms": round(ext_ms / max(launches, 1), 4),
        },
    }
    tr = measured_traffic() if args.workload == "icosphere" else None
    if tr is not None:
        # HBM bytes per ray from the profiled FETCH_SIZE + WRITE_SIZE, times this run's rays per
        # launch; the algorithmic bytes above count L2 / MALL hits too, so `frac` can pass 1
        R = line["roofline"]
        per_launch = tr[1] * R["rays_per_launch"]
        R["traffic"], R["traffic_per_ray"] = round(per_launch), round(tr[1], 1)
        if R["avg_launch_ms"] > 0:
            R["measured_hbm_gbs"] = round(per_launch / (R["avg_launch_ms"] / 1e3) / 1e9, 1)
            R["measured_hbm_frac"] = round(R["measured_hbm_gbs"] / HBM_PEAK_GBS, 4)
        R["traffic_source"] = os.path.relpath(TRAFFIC_FILE, ROOT)
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        try:
            line["cpu_baseline"] = cpu_baseline(scene, nodes, wl, threads)
        except Exception as e:  # the baseline is reported, never required
            line["cpu_baseline"] = {"value": None, "error": repr(e)}
    finite = np.isfinite(img[..., :3]).all(-1).mean()
    print(f"# image mean {img[..., :3].reshape(-1, 3).mean(0)}, finite {finite:.4f}, samples/pixel "
          f"{img[0, 0, 3]}", file=sys.stderr)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
