"""Benchmark of the path-tracing hot path on one MI355X.

Default run: one JSON line.  Its headline (`value`) is Mrays/s on the
~1M-triangle 1920x1080 scene (48 icospheres of 20,480 triangles in a
10x10x10 room, 2-triangle quad light), path tracing with NEE+MIS, max depth
8, Russian roulette from bounce 3.  A step is one batch of the wavefront
path tracer: 32 frames of 1920x1080 at 1 sample per pixel each (66.4M
paths) run from primary rays until every path has ended, then folded into
the progressive image; 2 steps are the 64 spp of the workload.  The same
line carries:

  roofline      the closest-hit trace's HBM roofline (measured HBM bytes,
                beside the algorithmic node + triangle bytes and the vector
                issue rate), and rows for the any-hit trace and the shade
                kernel, timed in a pass whose two traces do not overlap
  cpu_baseline  the CPU restatement of the reference on this host's cores
  configs       the Cornell box (512x512, Oren-Nayar walls, 1024 spp) and
                bidirectional (1280x720, 32 spp) workloads, each with the
                roofline rows of its kernels
  batch         64 jobs on 8 icosphere scenes (20k-2M triangles), random
                cameras, 128^2-1024^2, 16-256 spp, all submitted at once to
                the continuous-batching scheduler, launch configuration from
                the library's autotuner, against the plain per-job paths on
                the same jobs
  serving       the same job mix arriving as a Poisson stream, 20 jobs/s for
                60 s: jobs/s, p50/p99 latency, against the plain per-job path

--workload icosphere|cornell|bdpt|batch|serve runs one of them alone.

Rays counted on the GPU: every BVH query it performs (extension, NEE
shadow, MIS any-hit, BDPT connection).  An NEE ray the quad light already
blocks, and a MIS ray that misses the quad, need no query on the GPU; the
CPU restatement counts every traceRay call of the reference, those
included, so both sides also report samples per second, which count the
same thing.

HBM bytes per unit of work and the vector-issue rates come from the
rocprofv3 counter summaries committed under profiles/ (taken with this
file's own workloads and launch options); they are scaled by this run's
units per launch.

--gpus above 1 is refused: the benchmark runs on one GPU.
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
PROFILES = os.path.join(ROOT, "profiles")

WORKLOADS = {
    "icosphere": dict(
        # long traversals: the closest-hit and any-hit traces of an iteration overlap on two streams
        metric="Mrays/s at 1M tris 1080p", W=1920, H=1080, frames=32, steps=2, mode=0, trace_overlap=1,
        params=dict(mis=1, max_bounces=8, rr_threshold=2), summary="hbm_1m_1080p_pt.txt",
        desc="path tracing, 48 icospheres x 20480 tris + 12-tri room + 2-tri quad light ({n} tris), 1920x1080, "
             "{f} spp per step, max depth 8, RR from bounce 3, NEE+MIS"),
    "cornell": dict(
        metric="Mrays/s Cornell box 512x512", W=512, H=512, frames=256, steps=4, mode=0,
        params=dict(mis=1, max_bounces=8, rr_threshold=2, oren_nayar=1), summary="hbm_cornell_512_pt.txt",
        desc="path tracing, Cornell box ({n} tris, Oren-Nayar walls with sigma 0), 512x512, {f} spp per step, "
             "max depth 8, RR from bounce 3, NEE+MIS"),
    "bdpt": dict(
        metric="Mrays/s BDPT Cornell box 720p", W=1280, H=720, frames=8, steps=4, mode=1,
        params=dict(mis=1, bdpt_bounces=5), summary="hbm_cornell_720p_bdpt.txt",
        desc="bidirectional path tracing (bdpt.cl), Cornell box ({n} tris), 1280x720, {f} spp per step, "
             "5-vertex eye and light subpaths, all connections, NEE+MIS"),
}


def read_summary(name):
    """key: value lines of a committed profile summary (None if absent)."""
    vals = {}
    try:
        with open(os.path.join(PROFILES, name)) as f:
            for line in f:
                if ":" in line and not line.startswith(("#", " ")):
                    k, v = line.split(":", 1)
                    try:
                        vals[k.strip()] = float(v.strip())
                    except ValueError:
                        vals[k.strip()] = v.strip()
    except OSError:
        return None
    vals["source"] = name
    return vals


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


# ---------------------------------------------------------------- serving

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


def even_sample(n, k):
    """k indices spread evenly over range(n) (all of them for k <= 0)."""
    if k <= 0 or k >= n:
        return list(range(n))
    return sorted({int(i * n / k) for i in range(k)})


def serve_run(Y, np, serve, rate, duration, n_jobs, plain_jobs, frame_jobs, serve_mode=0, graph=0, trace_overlap=0,
              tune=True):
    """Served (through the scheduler) against the plain per-job paths on
    jobs of the same list, at equal spp and estimator (the images are
    identical).  serve False: all jobs submitted at t = 0 (batched).  tune:
    the library's autotuner picks the launch configuration first, on the
    list's first 512x512 job at twice the paths in flight (both sides then
    run with it)."""
    scenes = serving_scenes(Y, np)
    n_jobs = n_jobs or (int(duration * rate) if serve else 64)
    jobs = serving_jobs(Y, np, len(scenes), n_jobs, rate if serve else 0.0)
    lat, done_t = [], []
    with Y.Context(mem_budget=64 << 30, graph=graph, trace_overlap=trace_overlap, serve_mode=serve_mode) as ctx:
        t0 = time.perf_counter()
        sids = [ctx.scene(s) for s in scenes]
        print(f"# {len(scenes)} scenes, {sum(len(s.tris) for s in scenes)} triangles, built and uploaded in "
              f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
        tuned = None
        if tune:
            t0 = time.perf_counter()
            j = next((j for j in jobs if j["W"] == 512), jobs[0])
            frames = max(2, 2 * ctx.config()["batch_paths"] // (j["W"] * j["H"]))
            tuned = ctx.autotune(sids[j["scene"]], j["cam"], j["W"], j["H"], j["params"], frames=frames)
            print(f"# autotuned in {time.perf_counter() - t0:.1f}s: {tuned}", file=sys.stderr, flush=True)
        launch = {k: v for k, v in ctx.config().items()
                  if k in ("max_paths", "bvh_width", "lds_stack", "shade_waves", "trace_block", "trace_chunk",
                           "trace_xcd", "trace_refill", "graph", "trace_overlap", "batch_factor", "batch_paths",
                           "serve_mode")}
        # warm the kernels once
        w = ctx.job(sids[0], jobs[0]["cam"], 64, 64, jobs[0]["params"])
        w.render(1)
        w.release()
        out = {}
        keep = set(even_sample(len(jobs), 3))
        clock0 = ctx.clock_ms()
        nxt = 0
        live = {}
        last_print = time.perf_counter()
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
                    if k in keep:
                        out[k] = live[k].read()
                    live[k].release()
                    del live[k]
            if time.perf_counter() - last_print > 10:
                last_print = time.perf_counter()
                print(f"# {'serving' if serve else 'batch'}: {len(done_t)} of {len(jobs)} jobs done, "
                      f"{len(live)} in the scheduler", file=sys.stderr, flush=True)
        makespan = max(done_t) / 1e3
        sst = ctx.serve_stats()
        # plain paths: one job at a time (each job's frames batched), and the
        # reference's frame-at-a-time loop, on evenly spread samples of the list
        pj = even_sample(len(jobs), plain_jobs)
        fj = even_sample(len(jobs), frame_jobs)
        t_job = t_frame = 0.0
        same = True
        for k in pj:
            j = jobs[k]
            b = ctx.job(sids[j["scene"]], j["cam"], j["W"], j["H"], j["params"], base_seed=j["seed"])
            t1 = time.perf_counter()
            b.render(j["spp"])
            img = b.read()
            t_job += time.perf_counter() - t1
            if k in out:
                same &= bool(np.array_equal(img.view(np.int32), out[k].view(np.int32)))
            b.release()
        print(f"# plain job-at-a-time: {len(pj)} jobs in {t_job:.1f}s", file=sys.stderr, flush=True)
        for k in fj:
            j = jobs[k]
            a = ctx.job(sids[j["scene"]], j["cam"], j["W"], j["H"], j["params"], base_seed=j["seed"])
            t1 = time.perf_counter()
            for _ in range(j["spp"]):
                a.render(1)
            a.read()
            t_frame += time.perf_counter() - t1
            a.release()
        print(f"# plain frame-at-a-time: {len(fj)} jobs in {t_frame:.1f}s", file=sys.stderr, flush=True)
    lat = np.array(lat)
    served = len(jobs) / makespan
    samples = sum(j["spp"] * j["W"] * j["H"] for j in jobs)

    def list_rate(idx, t):
        """jobs/s of a plain path over the whole list: its time on the sampled
        jobs scaled by the list's samples over theirs (all jobs: exact)"""
        part = sum(jobs[k]["spp"] * jobs[k]["W"] * jobs[k]["H"] for k in idx)
        return len(jobs) / (t * samples / part)

    plain_job = list_rate(pj, t_job)
    plain_frame = list_rate(fj, t_frame)
    return {
        "value": round(served, 3), "unit": "jobs/s",
        "workload": ("Poisson arrivals at %.1f jobs/s for %.1f s" % (rate, duration) if serve else
                     "%d jobs submitted together" % len(jobs)) +
                    ", 8 icosphere scenes (20k-2M tris), random cameras, 128^2-1024^2, 16-256 spp, depth 8, "
                    "NEE+MIS, 64 GB image budget, %s%s" % ("continuous batching" if serve_mode == 0 else "rounds",
                                                          ", bounce loop replayed from a HIP graph"
                                                          if launch.get("graph") else ""),
        "jobs": len(jobs), "scene_tris": [len(s.tris) for s in scenes],
        "total_megasamples": round(samples / 1e6, 1),
        "launch": launch, "autotuned": tuned is not None,
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 1), "p99": round(float(np.percentile(lat, 99)), 1),
                       "max": round(float(lat.max()), 1)},
        "makespan_s": round(makespan, 3),
        "served_mrays_per_s": round((sst.rays_ext + sst.rays_shadow) / makespan / 1e6, 1),
        "served_msamples_per_s": round(samples / makespan / 1e6, 1),
        # kernel times are taken launch by launch; a replayed graph has none
        "served_kernels_ms": ({"trace_ext": round(sst.ms_trace_ext, 1), "trace_shadow": round(sst.ms_trace_shadow, 1),
                               "shade": round(sst.ms_shade, 1)} if not launch.get("graph") else None),
        "host_between_groups_ms": round(sst.ms_host, 1),
        "served_loop": {"groups": sst.groups, "refills": sst.refills, "plan_stalls": sst.plan_stalls,
                        "mean_paths_in_flight": round(sst.paths_in_flight / max(sst.groups, 1))},
        "plain": {"jobs_per_s_job_at_a_time": round(plain_job, 3), "job_at_a_time_jobs": len(pj),
                  "jobs_per_s_frame_at_a_time": round(plain_frame, 3), "frame_at_a_time_jobs": len(fj),
                  "how": "each plain path times the jobs named (spread evenly over the list); its jobs/s for the "
                         "whole list scales that time by the list's samples over theirs",
                  "speedup_vs_job_at_a_time": round(served / plain_job, 2),
                  "speedup_vs_frame_at_a_time": round(served / plain_frame, 2),
                  "images_identical_to_served": same},
    }


# ------------------------------------------------------- single workloads

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
            "samples_per_s": round(samples / dt, 1), "cpu_model": cpu_model(),
            "sample": f"{tw}x{th} centre tile of the same {W}x{H} frame, {spp} spp (frames 2..{1 + spp}), "
                      f"{samples} samples, {rays} traceRay calls in {dt:.1f} s, {threads} threads"}


def context_kwargs(args, wl, trace_overlap=None):
    ov = trace_overlap if trace_overlap is not None else (
        args.trace_overlap if args.trace_overlap >= 0 else wl.get("trace_overlap", 0))
    return dict(bvh_width=args.bvh_width, trace_block=args.trace_block, trace_chunk=args.trace_chunk,
                leaf_max=args.leaf_max, hoist_frac=args.hoist_frac, trace_grid=args.trace_grid,
                lds_stack=args.lds_stack, shade_waves=args.shade_waves, max_paths=args.max_paths,
                batch_factor=args.batch_factor, trace_overlap=ov, graph=max(args.graph, 0),
                trace_xcd=args.trace_xcd or wl.get("trace_xcd", 0),
                trace_refill=args.trace_refill or wl.get("trace_refill", 0), bvh_build=args.bvh_build)


STAT_KEYS = ("rays_ext", "rays_shadow", "samples", "ms_trace_ext", "ms_trace_shadow", "ms_shade",
             "trace_launches_ext", "trace_launches_shadow", "iterations")


def pt_run(Y, np, name, args, steps, warmup, count=True, trace_overlap=None, scene=None, nodes=None):
    """Timed steps of one single-job workload: the stats of the timed steps,
    the counting pass's stats and the scene."""
    wl = WORKLOADS[name]
    W, H, F = wl["W"], wl["H"], args.frames or wl["frames"]
    if scene is None:
        t0 = time.perf_counter()
        scene = make_scene(Y, name)
        nodes = scene.bvh()
        print(f"# {name}: scene {len(scene.tris)} triangles, {len(nodes)} BVH nodes, build "
              f"{time.perf_counter() - t0:.2f}s", file=sys.stderr, flush=True)
    params = Y.params_reference(wl["mode"], **wl["params"])
    with Y.Context(**context_kwargs(args, wl, trace_overlap)) as ctx:
        sid = ctx.scene(scene, nodes)
        info = ctx.scene_info(sid)
        launch = {k: v for k, v in ctx.config().items()
                  if k in ("max_paths", "lds_stack", "shade_waves", "trace_block", "trace_chunk", "leaf_max",
                           "trace_xcd", "trace_refill", "graph", "trace_overlap", "batch_factor", "bvh_build")}
        cst = None
        if count:   # counting pass: exact node visits / triangle tests of one step
            cj = ctx.job(sid, scene.cam, W, H, params, base_seed=BASE_SEED)
            cj.set_counting(True)
            cj.render(F)
            cst = cj.stats()
            cj.release()
        job = ctx.job(sid, scene.cam, W, H, params, base_seed=BASE_SEED)
        for _ in range(warmup):
            job.render(F)
        s0 = job.stats()
        t0 = time.perf_counter()
        for _ in range(steps):
            job.render(F)
        dt = time.perf_counter() - t0
        s1 = job.stats()
        img = job.read()
    d = {k: getattr(s1, k) - getattr(s0, k) for k in STAT_KEYS}
    return dict(wl=wl, name=name, F=F, steps=steps, warmup=warmup, dt=dt, d=d, cst=cst, info=info, launch=launch,
                scene=scene, nodes=nodes, img=img)


def kernel_row(kernel, work_unit, units, launches, ms, summary, key, alg_bytes=None, valu_key=None):
    """Roofline row of one kernel: units of work per launch and the launch
    duration from this run (HIP events on the kernel's stream), HBM bytes
    per unit and the vector-issue rate from the committed counter summary."""
    launches = max(launches, 1)
    per_launch = units / launches
    t = ms / launches
    row = {"kernel": kernel, "work_unit": work_unit, "units_per_launch": round(per_launch),
           "avg_launch_ms": round(t, 4)}
    hb = summary.get(key) if summary else None
    if isinstance(hb, float) and t > 0:
        gbs = hb * per_launch / (t / 1e3) / 1e9
        row.update({"bound": "hbm", "hbm_bytes_per_unit": round(hb, 1), "traffic": round(hb * per_launch),
                    "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(gbs / HBM_PEAK_GBS, 4)})
    if alg_bytes is not None and t > 0:
        ag = alg_bytes * per_launch / (t / 1e3) / 1e9
        row["algorithmic"] = {"bytes_per_unit": round(alg_bytes, 1), "gbs": round(ag, 1),
                              "per_hbm_peak": round(ag / HBM_PEAK_GBS, 4)}
    if summary and valu_key:
        vb, vu = summary.get(valu_key + "_valu_busy"), summary.get(valu_key + "_valu_util")
        if isinstance(vb, float):
            row["valu_busy"] = round(vb / 100.0, 3)
        if isinstance(vu, float):
            row["valu_util"] = round(vu / 100.0, 3)
        # wave-level instructions per unit of work and the texture address /
        # data units' busy share (the memory pipeline in front of the L1)
        u = work_unit.split()[-1]
        for part in ("valu", "salu", "vmem"):
            x = summary.get(f"{valu_key}_{part}_insts_per_{u}")
            if isinstance(x, float):
                row[f"{part}_insts_per_{u}"] = round(x, 2)
        for part in ("ta", "td"):
            x = summary.get(f"{valu_key}_{part}_busy")
            if isinstance(x, float):
                row[f"{part}_busy"] = round(x / 100.0, 3)
    if summary:
        row["counters_source"] = "profiles/" + summary["source"]
    return row


def rows_for(res, summary):
    """Roofline rows of a workload's kernels from its timed steps."""
    d, cst, info = res["d"], res["cst"], res["info"]
    node_b = 64 if info.width == 2 else 128
    bdpt = res["wl"]["mode"] == 1
    alg_ext = alg_sh = None
    if cst is not None:
        # algorithmic bytes per ray from the counting pass: interior nodes
        # visited x node size + triangle tests x 48 B + the ray (32 B) and its
        # result (closest hit 16 B, any hit 4 B)
        alg_ext = (cst.nodes_ext * node_b + cst.tris_ext * 48) / max(cst.rays_ext, 1) + 32 + 16
        alg_sh = (cst.nodes_shadow * node_b + cst.tris_shadow * 48) / max(cst.rays_shadow, 1) + 32 + 4
    return [
        kernel_row("k_trace_rf closest hit", "ray", d["rays_ext"], d["trace_launches_ext"], d["ms_trace_ext"],
                   summary, "trace_ext_bytes_per_ray", alg_ext, "trace_ext"),
        kernel_row("k_trace_rf any hit", "ray", d["rays_shadow"], d["trace_launches_shadow"], d["ms_trace_shadow"],
                   summary, "trace_any_bytes_per_ray", alg_sh, "trace_any"),
        # every ray traced comes back to the shade pass as one path event
        kernel_row("k_shade_bdpt" if bdpt else "k_shade_pt", "path event", d["rays_ext"] + d["rays_shadow"],
                   d["iterations"], d["ms_shade"], summary, "shade_bytes_per_event", None, "shade"),
    ]


def config_summary(res):
    d, wl, F = res["d"], res["wl"], res["F"]
    rays = d["rays_ext"] + d["rays_shadow"]
    return {
        "workload": wl["desc"].format(n=len(res["scene"].tris), f=F),
        "triangles": len(res["scene"].tris), "bvh_nodes_ref": len(res["nodes"]), "bvh_width": res["info"].width,
        "resolution": [wl["W"], wl["H"]], "spp_per_step": F, "spp_total": F * res["steps"],
        "rays": int(rays), "rays_per_sample": round(rays / max(d["samples"], 1), 3),
        "samples_per_s": round(d["samples"] / res["dt"], 1),
        "ms_trace_ext": round(d["ms_trace_ext"], 2), "ms_trace_shadow": round(d["ms_trace_shadow"], 2),
        "ms_shade": round(d["ms_shade"], 2), "launch": res["launch"],
    }


def workload_line(Y, np, name, args, with_cpu):
    """A single workload's JSON object (the headline's form)."""
    wl = WORKLOADS[name]
    steps = args.steps or wl["steps"]
    res = pt_run(Y, np, name, args, steps, args.warmup, count=not args.no_count)
    summary = read_summary(wl["summary"])
    rows = rows_for(res, summary)
    if wl.get("trace_overlap", 0) and args.trace_overlap != 0:
        # the timed steps overlap the two traces; the kernels' own durations
        # come from one more step with the traces one after the other
        rr = pt_run(Y, np, name, args, 1, 1, count=False, trace_overlap=0, scene=res["scene"], nodes=res["nodes"])
        rr["cst"] = res["cst"]
        rows = rows_for(rr, summary)
        for r in rows:
            r["timing"] = "one step with the two traces not overlapped"
    d = res["d"]
    rays = d["rays_ext"] + d["rays_shadow"]
    line = {
        "metric": wl["metric"], "value": round(rays / res["dt"] / 1e6, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": round(res["dt"] * 1e3 / max(steps, 1), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": config_summary(res),
    }
    # the roofline of the kernel with the longest launches: the closest-hit
    # trace on the 1M-triangle scene, the shade pass on the Cornell box
    main_row = max(rows, key=lambda r: r["avg_launch_ms"])
    R = {"kernel": main_row["kernel"]}
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        R[k] = main_row.get(k)
    if R["bound"] is None:
        R.update({"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s"})
    if "algorithmic" in main_row:
        R["algorithmic"] = dict(main_row["algorithmic"])
        if res["cst"] is not None:
            c = res["cst"]
            R["algorithmic"]["nodes_per_ray"] = round(c.nodes_ext / max(c.rays_ext, 1), 2)
            R["algorithmic"]["tri_tests_per_ray"] = round(c.tris_ext / max(c.rays_ext, 1), 2)
    if "valu_busy" in main_row:
        R["valu_busy"] = main_row["valu_busy"]
        R["valu_util"] = main_row.get("valu_util")
        for k in ("td_busy", "ta_busy"):
            if k in main_row:
                R[k] = main_row[k]
        # what limits the kernel: the largest of its HBM fraction, its
        # vector-issue fraction and the busy share of the texture data
        # unit (the vector-memory pipeline's return path into registers)
        cands = {"hbm": R["frac"] or 0.0, "vector issue": main_row["valu_busy"],
                 "vector memory pipeline": main_row.get("td_busy", 0.0)}
        R["limiter"] = max(cands, key=cands.get)
        R["bound_frac"] = round(max(cands.values()), 4)
    R["units_per_launch"] = main_row["units_per_launch"]
    R["avg_launch_ms"] = main_row["avg_launch_ms"]
    R["rows"] = rows
    line["roofline"] = R
    if with_cpu:
        threads = min(16, os.cpu_count() or 1)
        try:
            cb = cpu_baseline(res["scene"], res["nodes"], wl, threads)
            cb["gpu_samples_per_s"] = round(d["samples"] / res["dt"], 1)
            line["cpu_baseline"] = cb
        except Exception as e:  # the baseline is reported, never required
            line["cpu_baseline"] = {"value": None, "error": repr(e)}
    img = res["img"]
    finite = np.isfinite(img[..., :3]).all(-1).mean()
    print(f"# {name}: {line['value']} Mrays/s, image mean {img[..., :3].reshape(-1, 3).mean(0)}, finite "
          f"{finite:.4f}, samples/pixel {img[0, 0, 3]}", file=sys.stderr, flush=True)
    return line


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed steps of the headline (0: the workload's default)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["all"] + sorted(WORKLOADS) + ["batch", "serve"], default="all",
                    help="all (default): the headline line with every other workload in it")
    ap.add_argument("--rate", type=float, default=20.0, help="serve: Poisson arrival rate (jobs/s)")
    ap.add_argument("--duration", type=float, default=60.0, help="serve: seconds of arrivals")
    ap.add_argument("--jobs", type=int, default=0, help="batch / serve: number of jobs (0: 64 / rate x duration)")
    ap.add_argument("--plain-jobs", type=int, default=-1,
                    help="jobs timed on the plain job-at-a-time path, spread over the list (0: all; -1: all for "
                         "batch, 120 for serve)")
    ap.add_argument("--frame-jobs", type=int, default=16,
                    help="jobs timed on the frame-at-a-time loop, spread over the list (0: all)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-count", action="store_true",
                    help="skip the counting pass (node visits, triangle tests): for counter runs, whose kernels "
                         "then belong to the timed steps only")
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
    ap.add_argument("--graph", type=int, default=-1,
                    help="1: bounce loop replayed from a HIP graph (-1: on for batch / serve, off for the workloads "
                         "whose roofline rows need the kernels timed one by one)")
    ap.add_argument("--trace-xcd", type=int, default=0, help="ray-queue slices per trace launch (1: no slicing)")
    ap.add_argument("--trace-refill", type=int, default=0, help="idle lanes before a trace wave takes new rays (-1: off)")
    ap.add_argument("--leaf-max", type=int, default=0, help="triangles per BVH leaf (0: library default)")
    ap.add_argument("--bvh-build", type=int, default=0, help="GPU tree: 1 own SAH build, 2 the reference's tree")
    ap.add_argument("--hoist-frac", type=float, default=0.0, help="box-area fraction above which triangles are hoisted")
    ap.add_argument("--serve-mode", type=int, default=0, help="batch / serve: 0 continuous batching, 1 rounds")
    ap.add_argument("--no-tune", action="store_true", help="batch / serve: library defaults instead of the autotuner")
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

    def serving(kind):
        serve = kind == "serve"
        plain = args.plain_jobs if args.plain_jobs >= 0 else (120 if serve else 0)
        return serve_run(Y, np, serve, args.rate, args.duration, args.jobs, plain, args.frame_jobs,
                         serve_mode=args.serve_mode, graph=1 if args.graph < 0 else args.graph,
                         trace_overlap=max(args.trace_overlap, 0),
                         tune=not args.no_tune)

    if args.workload in ("batch", "serve"):
        r = serving(args.workload)
        line = {"metric": "served jobs/s with p50/p99 latency vs per-job" if args.workload == "serve"
                else "batched jobs/s vs per-job", "value": r.pop("value"), "unit": r.pop("unit"), "n_gpus": 1,
                "steps": r["jobs"], "warmup": 1, "ms_per_step": round(r["makespan_s"] * 1e3 / r["jobs"], 3),
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                "data": "synthetic", "config": {"workload": r.pop("workload")}}
        line.update(r)
        print(json.dumps(line))
        return
    if args.workload != "all":
        print(json.dumps(workload_line(Y, np, args.workload, args, not args.no_cpu)))
        return
    line = workload_line(Y, np, "icosphere", args, not args.no_cpu)
    sub = argparse.Namespace(**vars(args))
    sub.steps, sub.frames, sub.trace_overlap = 0, 0, -1
    line["configs"] = {}
    for name in ("cornell", "bdpt"):
        r = workload_line(Y, np, name, sub, False)
        line["configs"][name] = {"metric": r["metric"], "value": r["value"], "unit": r["unit"], "steps": r["steps"],
                                 "ms_per_step": r["ms_per_step"], "config": r["config"], "roofline": r["roofline"]}
    line["batch"] = serving("batch")
    line["serving"] = serving("serve")
    print(json.dumps(line))


if __name__ == "__main__":
    main()
