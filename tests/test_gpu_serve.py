"""GPU tests of the kernel-contract entry point and of the scheduler.

* yn_pathtracer (the reference kernel's 14 arguments: tiles, reset, rand,
  ping-pong images) gives the same image as yn_job_render with the same
  frame seeds, bit for bit;
* jobs served together (several scenes and sizes in one wavefront batch)
  give the same image as each job rendered alone, bit for bit: a path's
  arithmetic depends only on its pixel, frame seed and scene, and each
  pixel's frames are folded in frame order either way;
* admission against the memory budget holds jobs back until images are
  released, and they still finish with the same result;
* continuous batching: jobs submitted while others are on the GPU join the
  running wavefront as paths end (they start before the running jobs
  finish), frames larger than the paths in flight enter in parts, and
  every image is still the job's image alone;
* the batched mix of the benchmark (64 jobs on 8 icosphere scenes of
  20k-2M triangles, the benchmark's own job list with 1/16 of its spp):
  every served image equals the job alone, and tiles of three jobs equal
  the oracle's.
"""
import numpy as np
import pytest

import yune_amd as Y
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def test_kernel_contract_tiles_and_ping_pong_match_job():
    import torch
    s = Y.gen_cornell()
    W, H = 70, 50                     # tiles of uneven size (ceil(W/2), ceil(H/2))
    p = Y.params_reference(mis=1, max_bounces=8, rr_threshold=2)
    seeds = O.mt19937(9, 2)
    with Y.Context(max_paths=1 << 16) as c:
        sid = c.scene(s)
        a = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        b = torch.zeros_like(a)
        for blk in range(4):          # frame 0: reset, into a
            c.pathtracer(a, None, s.cam, sid, W, H, p, 1, 1, int(seeds[0]), blk, 2, 2)
        for blk in range(4):          # frame 1: read a, write b (ping-pong)
            c.pathtracer(b, a, s.cam, sid, W, H, p, 1, 0, int(seeds[1]), blk, 2, 2)
        torch.cuda.synchronize()
        got = b.cpu().numpy()
        job = c.job(sid, s.cam, W, H, p, base_seed=9)
        job.render(2)
        ref = job.read()
    assert np.array_equal(got.view(np.int32), ref.view(np.int32))
    assert (got[..., 3] == 2).all()


def _jobs(scenes):
    rng = np.random.default_rng(4)
    out = []
    for k in range(7):
        sid = int(rng.integers(len(scenes)))
        W, H = [(48, 32), (64, 64), (33, 21), (96, 40)][k % 4]
        mode = Y.YN_MODE_BDPT if k == 5 else Y.YN_MODE_PT
        kw = dict(bdpt_bounces=4) if mode else dict(mis=k % 2, max_bounces=8, rr_threshold=2)
        out.append(dict(sid=sid, W=W, H=H, p=Y.params_reference(mode, **kw), frames=int(rng.integers(2, 9)),
                        seed=int(rng.integers(1, 1 << 30))))
    return out


@pytest.mark.parametrize("serve_mode,graph", [(0, 0), (1, 0), (0, 1)])
def test_served_jobs_equal_jobs_alone(serve_mode, graph):
    """serve_mode 0: continuous batching (the bidirectional job runs in a
    round of its own); 1: scheduler rounds; graph 1: the bounce loop
    replayed from a captured graph."""
    scenes = [Y.gen_cornell(), Y.gen_icosphere_room(3, 9, 2)]
    with Y.Context(max_paths=1 << 15, serve_mode=serve_mode, graph=graph) as c:
        sids = [c.scene(s) for s in scenes]
        specs = _jobs(scenes)
        jobs = [c.submit(sids[j["sid"]], scenes[j["sid"]].cam, j["W"], j["H"], j["p"], j["frames"], j["seed"])
                for j in specs]
        for _ in range(1000):
            if c.serve(4) == 0:
                break
        served = []
        for jb, j in zip(jobs, specs):
            st = jb.status()
            assert st["done"] == 1 and st["frames_done"] == j["frames"]
            assert st["t_done_ms"] >= st["t_start_ms"] >= st["t_submit_ms"] >= 0
            served.append(jb.read())
            jb.release()
        for img, j in zip(served, specs):
            alone = c.job(sids[j["sid"]], scenes[j["sid"]].cam, j["W"], j["H"], j["p"], base_seed=j["seed"])
            alone.render(j["frames"])
            ref = alone.read()
            alone.release()
            assert np.array_equal(img.view(np.int32), ref.view(np.int32))
        assert c.serve_stats().rays_ext > 0


@pytest.mark.parametrize("mode", ["pt", "bdpt"])
def test_material_table_follows_scenes(mode):
    """Launches over several scenes read every live scene's materials (and
    light quads) from tables staged in LDS, each scene's materials at its
    own offset; releasing a scene and making another rebuilds the table and
    moves the offsets.  Served (several scenes per launch) equals alone (one
    scene per launch, its own materials) bit for bit after each change."""
    import dataclasses
    cornell, room = Y.gen_cornell(), Y.gen_icosphere_room(3, 9, 2)
    # the Cornell box with its material list doubled: same image, more entries
    wide = dataclasses.replace(cornell, mats=np.concatenate([cornell.mats, cornell.mats]), nodes=None)
    p = (Y.params_reference(mis=1, max_bounces=8, rr_threshold=2) if mode == "pt" else
         Y.params_reference(Y.YN_MODE_BDPT, bdpt_bounces=4))

    def served_equals_alone(c, pairs):
        jobs = [c.submit(sid, s.cam, 40, 30, p, 3, 11 + k) for k, (sid, s) in enumerate(pairs)]
        for _ in range(1000):
            if c.serve(4) == 0:
                break
        for k, ((sid, s), jb) in enumerate(zip(pairs, jobs)):
            img = jb.read()
            jb.release()
            alone = c.job(sid, s.cam, 40, 30, p, base_seed=11 + k)
            alone.render(3)
            ref = alone.read()
            alone.release()
            assert np.array_equal(img.view(np.int32), ref.view(np.int32)), k

    with Y.Context(max_paths=1 << 15) as c:
        a = c.scene(wide)              # slot 0: the room's offset is 2n
        b = c.scene(room)
        served_equals_alone(c, [(a, wide), (b, room)])
        c.release_scene(a)
        d = c.scene(cornell)           # slot 0 again: the room's offset is n
        served_equals_alone(c, [(d, cornell), (b, room)])


def test_admission_waits_for_memory():
    s = Y.gen_cornell()
    W, H = 64, 64
    img_bytes = W * H * 16
    p = Y.params_reference(mis=1, max_bounces=8, rr_threshold=2)
    with Y.Context(max_paths=1 << 15, mem_budget=2 * img_bytes + 100) as c:
        sid = c.scene(s)
        jobs = [c.submit(sid, s.cam, W, H, p, 3, base_seed=10 + k) for k in range(4)]
        st = [j.status() for j in jobs]
        assert [x["admitted"] for x in st] == [1, 1, 0, 0]
        with pytest.raises(Y.YuneError):
            jobs[3].read()
        done = {}
        for _ in range(200):
            c.serve(1)
            for k, j in enumerate(jobs):
                if k not in done and j.status()["done"]:
                    done[k] = j.read()
                    j.release()
            if len(done) == 4:
                break
        assert len(done) == 4
        for k in range(4):
            alone = c.job(sid, s.cam, W, H, p, base_seed=10 + k)
            alone.render(3)
            assert np.array_equal(done[k].view(np.int32), alone.read().view(np.int32))
            alone.release()


def _read_png(path):
    import struct
    import zlib
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat = 8, b""
    w = h = None
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert zlib.crc32(typ + body) == crc
        if typ == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert (depth, ctype) == (8, 2)
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
    assert (rows[:, 0] == 0).all()                 # filter type 0
    return rows[:, 1:].reshape(h, w, 3)


def _read_hdr(path):
    data = open(path, "rb").read()
    head, _, rest = data.partition(b"\n\n")
    assert head.startswith(b"#?RADIANCE") and b"FORMAT=32-bit_rle_rgbe" in head
    res, _, px = rest.partition(b"\n")
    _, h, _, w = res.split()
    h, w = int(h), int(w)
    e = np.frombuffer(px, np.uint8).reshape(h, w, 4).astype(np.float64)
    scale = np.where(e[..., 3:] > 0, np.ldexp(1.0, (e[..., 3:] - 136).astype(int)), 0.0)
    return (e[..., :3] + 0.5) * scale * (e[..., 3:] > 0)


def test_save_hdr_png_flip_and_tonemap(tmp_path):
    """RendererCore::saveImage: the file's first row is the image's last
    (vertical flip), .hdr holds the linear image, .png the tonemapped one
    as round(clamp(v) * 255); the device tonemap equals the oracle's."""
    s = Y.gen_cornell()
    W, H = 40, 24
    with Y.Context(max_paths=1 << 14) as c:
        job = c.job(c.scene(s), s.cam, W, H, Y.params_reference(mis=1, max_bounces=8, rr_threshold=2), base_seed=3)
        job.render(4)
        img = job.read()
        tm = job.read_tonemapped()
        job.save(str(tmp_path / "a.hdr"))
        job.save(str(tmp_path / "a.png"))
    assert np.array_equal(tm.view(np.int32), O.tonemap(img).view(np.int32))
    # the Hable variant on the device equals the oracle's, bit for bit, both
    # through the job (its .png) and on caller-owned device images
    with Y.Context(max_paths=1 << 14) as c:
        job = c.job(c.scene(s), s.cam, W, H, Y.params_reference(mis=1, max_bounces=8, rr_threshold=2), base_seed=3)
        job.render(4)
        job.set_tonemap(Y.TONEMAP_HABLE)
        th = job.read_tonemapped()
        job.save(str(tmp_path / "h.png"))
        import torch
        src = torch.from_numpy(img).cuda()
        dst = torch.empty_like(src)
        c.tonemap(dst, src, W, H, Y.TONEMAP_HABLE)
        td = dst.cpu().numpy()
        with pytest.raises(Y.YuneError):
            job.set_tonemap(2)
    oh = O.tonemap(img, 1)
    assert np.array_equal(th.view(np.int32), oh.view(np.int32))
    assert np.array_equal(td.view(np.int32), oh.view(np.int32))
    hp = _read_png(tmp_path / "h.png")
    exp = np.rint(np.clip(np.nan_to_num(oh[..., :3], nan=0.0), 0, 1) * 255).astype(np.uint8)[::-1]
    assert np.abs(hp.astype(int) - exp.astype(int)).max() <= 1
    png = _read_png(tmp_path / "a.png")
    exp = np.rint(np.clip(np.nan_to_num(tm[..., :3], nan=0.0), 0, 1) * 255).astype(np.uint8)[::-1]
    assert np.abs(png.astype(int) - exp.astype(int)).max() <= 1    # lround vs rint on exact halves
    hdr = _read_hdr(tmp_path / "a.hdr")
    ref = img[..., :3][::-1].astype(np.float64)
    # RGBE: one shared exponent per pixel, 8-bit mantissas
    assert (np.abs(hdr - ref) <= ref.max(-1, keepdims=True) / 128 + 1e-6).all()


def test_autotune_chooses_valid_config_and_keeps_results():
    """yn_autotune tries every knob (BVH width rebuilds the scene layout) and
    leaves a configuration under which images stay bit-identical to the
    oracle."""
    s = Y.gen_icosphere_room(3, 9, 2)
    W, H = 80, 45
    kw = dict(mis=1, max_bounces=8, rr_threshold=2)
    p = Y.params_reference(**kw)
    with Y.Context(max_paths=1 << 16) as c:
        sid = c.scene(s)
        r = c.autotune(sid, s.cam, W, H, p, frames=2)
        assert r["n_tried"] >= 12 and r["mrays_per_s"] > 0
        assert r["bvh_width"] in (2, 4) and r["trace_block"] in (64, 128, 256)
        assert r["lds_stack"] in (0, 8, 12, 16, 32) and r["shade_waves"] in (2, 3, 4)
        cfg = c.config()
        assert (cfg["bvh_width"], cfg["trace_chunk"], cfg["batch_paths"]) == (r["bvh_width"], r["trace_chunk"],
                                                                               r["batch_paths"])
        assert c.scene_info(sid).width == r["bvh_width"]
        job = c.job(sid, s.cam, W, H, p, base_seed=11)
        job.render(4)
        img = job.read()
    oimg, _, _, _ = O.OScene(s.tris, s.mats, s.bvh(), s.light, heap_size=0).render(
        O.Cam.from_buffer_copy(bytes(s.cam)), W, H, O.params(0, **kw), O.mt19937(11, 4), threads=8)
    assert (img.view(np.int32) == oimg.view(np.int32)).all(-1).mean() >= 0.999


def test_arrivals_join_running_work():
    """Jobs submitted while two larger ones are in flight enter the
    running wavefront before those finish, and every image equals the job
    rendered alone."""
    s = Y.gen_cornell()
    room = Y.gen_icosphere_room(5, 6, 2)
    p = Y.params_reference(mis=1, max_bounces=8, rr_threshold=2)
    with Y.Context(max_paths=1 << 14, serve_mode=0) as c:
        sids = [c.scene(s), c.scene(room)]
        cams = [s.cam, room.cam]
        specs = [dict(sc=0, W=128, H=96, f=6, seed=21), dict(sc=1, W=96, H=96, f=5, seed=22)]
        jobs = [c.submit(sids[j["sc"]], cams[j["sc"]], j["W"], j["H"], p, j["f"], j["seed"]) for j in specs]
        c.serve(1)
        st = [j.status() for j in jobs]
        assert all(x["frames_done"] < x["frames_target"] for x in st), st
        late = [dict(sc=1, W=40, H=30, f=3, seed=31), dict(sc=0, W=64, H=64, f=2, seed=32),
                dict(sc=0, W=17, H=9, f=4, seed=33)]
        t_late = c.clock_ms()
        jobs += [c.submit(sids[j["sc"]], cams[j["sc"]], j["W"], j["H"], p, j["f"], j["seed"]) for j in late]
        specs += late
        for _ in range(2000):
            if c.serve(1) == 0:
                break
        st = [j.status() for j in jobs]
        assert all(x["done"] for x in st)
        # jobs run in submission order: the late ones enter while the
        # running jobs' last frames are still in flight
        last_done = max(x["t_done_ms"] for x in st[:2])
        for x in st[2:]:
            assert t_late <= x["t_start_ms"] < last_done, (x, last_done)
        imgs = [j.read() for j in jobs]
        for j in jobs:
            j.release()
        for img, j in zip(imgs, specs):
            alone = c.job(sids[j["sc"]], cams[j["sc"]], j["W"], j["H"], p, base_seed=j["seed"])
            alone.render(j["f"])
            assert np.array_equal(img.view(np.int32), alone.read().view(np.int32))
            alone.release()


def test_frames_larger_than_the_paths_in_flight():
    """A frame of 6,400 pixels with 4,096 paths in flight enters the
    continuous loop in parts (segments starting mid-frame)."""
    s = Y.gen_cornell()
    p = Y.params_reference(mis=1, max_bounces=8, rr_threshold=2)
    with Y.Context(max_paths=1 << 12, serve_mode=0) as c:
        sid = c.scene(s)
        jobs = [c.submit(sid, s.cam, 80, 80, p, 3, 41), c.submit(sid, s.cam, 50, 30, p, 2, 42)]
        for _ in range(2000):
            if c.serve(1) == 0:
                break
        imgs = [j.read() for j in jobs]
        for j in jobs:
            j.release()
        for img, (W, H, f, seed) in zip(imgs, [(80, 80, 3, 41), (50, 30, 2, 42)]):
            alone = c.job(sid, s.cam, W, H, p, base_seed=seed)
            alone.render(f)
            assert np.array_equal(img.view(np.int32), alone.read().view(np.int32))
            alone.release()


def test_releasing_a_job_in_flight():
    """A served job released while its paths are in flight: the others
    still finish with their own images."""
    s = Y.gen_cornell()
    p = Y.params_reference(mis=1, max_bounces=8, rr_threshold=2)
    with Y.Context(max_paths=1 << 13, serve_mode=0) as c:
        sid = c.scene(s)
        a = c.submit(sid, s.cam, 96, 96, p, 8, 51)
        b = c.submit(sid, s.cam, 40, 40, p, 3, 52)
        c.serve(1)
        a.release()
        for _ in range(2000):
            if c.serve(1) == 0:
                break
        assert b.status()["done"]
        img = b.read()
        b.release()
        alone = c.job(sid, s.cam, 40, 40, p, base_seed=52)
        alone.render(3)
        assert np.array_equal(img.view(np.int32), alone.read().view(np.int32))


def test_batched_mix_matches_alone_and_oracle():
    import bench
    scenes = bench.serving_scenes(Y, np)
    specs = bench.serving_jobs(Y, np, len(scenes), 64, 0.0)
    for j in specs:
        j["spp"] = max(1, j["spp"] // 16)
    print(f"\nmix: {len(specs)} jobs, scenes of {[len(x.tris) for x in scenes]} triangles, "
          f"{sum(j['spp'] * j['W'] * j['H'] for j in specs) / 1e6:.1f}M samples", flush=True)
    with Y.Context(mem_budget=64 << 30, serve_mode=0) as c:
        sids = [c.scene(x) for x in scenes]
        jobs = [c.submit(sids[j["scene"]], j["cam"], j["W"], j["H"], j["params"], j["spp"], j["seed"]) for j in specs]
        for _ in range(100000):
            if c.serve(4) == 0:
                break
        served = []
        for jb in jobs:
            assert jb.status()["done"]
            served.append(jb.read())
            jb.release()
        for img, j in zip(served, specs):
            alone = c.job(sids[j["scene"]], j["cam"], j["W"], j["H"], j["params"], base_seed=j["seed"])
            alone.render(j["spp"])
            assert np.array_equal(img.view(np.int32), alone.read().view(np.int32)), j
            alone.release()
    # tiles of three jobs on the three smallest scenes against the oracle
    small = np.argsort([len(x.tris) for x in scenes])[:3]
    for sc in small:
        k = next(i for i, j in enumerate(specs) if j["scene"] == sc)
        j = specs[k]
        x = scenes[sc]
        T = 24
        x0, y0 = j["W"] // 2 - T // 2, j["H"] // 2 - T // 2
        osc = O.OScene(x.tris, x.mats, x.bvh(), x.light, heap_size=0)
        oimg, _, _, _ = osc.render(O.Cam.from_buffer_copy(bytes(j["cam"])), j["W"], j["H"], O.params(0, mis=1, max_bounces=8,
                                   rr_threshold=2), O.mt19937(j["seed"], j["spp"]), rect=(x0, y0, x0 + T, y0 + T),
                                   threads=8)
        g = served[k][y0:y0 + T, x0:x0 + T]
        o = oimg[y0:y0 + T, x0:x0 + T]
        same = (g.view(np.int32) == o.view(np.int32)).all(-1).mean()
        a, b = g[..., :3].astype(np.float64), o[..., :3].astype(np.float64)
        err = float((np.abs(a - b).sum(-1) / np.maximum(np.abs(b).sum(-1), 1e-2)).mean())
        print(f"  job {k} (scene {sc}, {len(x.tris)} triangles): tile bit-identical {same:.4f}, mean rel err {err:.2e}")
        assert same >= 0.99 and err <= 0.01, (k, same, err)
