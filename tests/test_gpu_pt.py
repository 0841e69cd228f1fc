"""GPU parity of the HIP path tracer against the CPU oracle.

The three parity bars (closest-hit ids and t, first-shadow-ray results, and
per-pixel radiance) are measured on the same scenes, cameras and frame
seeds on both sides.  The oracle runs here with an uncapped traversal list
(heap_size = 0, the exact closest hit) except where a test says otherwise;
the full-size test runs it with the reference's HEAP_SIZE = 1500 list and
checks that no query there reached the cap.

Tolerances:
  * closest hit: same triangle id, and the same t bit for bit, except for
    rays whose two candidate triangles are at equal t within 1e-5 relative
    (a tie, resolved by test order), which are counted and must be rare;
  * shadow rays: same occluded flag, and the GPU's occluder must be a real
    occluder according to the oracle;
  * radiance: bit-identical in at least 99.9% of pixels, and mean over
    pixels of |gpu - cpu|_1 / max(|cpu|_1, 1e-2) <= 1%.  The kernels and
    the oracle share one definition of sin, cos and pow (yn_fmath.h) and
    every other float operation rounds once in the reference's order, so
    only closest-hit ties (equal t on two triangles) can make paths differ.
"""
import os
import time

import numpy as np
import pytest

import yune_amd as Y
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def oscene(s, nodes, heap=0):
    return O.OScene(s.tris, s.mats, nodes, s.light, heap_size=heap)


def ocam(cam):
    return O.Cam.from_buffer_copy(bytes(cam))


def random_rays(n, seed, lo, hi, shadow_frac=0.3):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 4:7] = d.astype(np.float32)
    shadow = rng.uniform(size=n) < shadow_frac
    rays[:, 3] = np.where(shadow, rng.uniform(0.05, 6.0, size=n), np.inf).astype(np.float32)
    rays[:, 7] = shadow.astype(np.float32)
    return rays


def tri_t(s, tri_id, ray):
    t = np.array([ray[3] if np.isfinite(ray[3]) else np.inf], np.float32)
    tri = np.ascontiguousarray(s.tris[tri_id:tri_id + 1])
    o = np.array([ray[0], ray[1], ray[2], 1], np.float32)
    d = np.array([ray[4], ray[5], ray[6], 0], np.float32)
    import ctypes as C
    L = C.c_float(np.inf)
    got = O.lib().or_ray_tri(O._p(tri), O._p(o), O._p(d), C.byref(L))
    return got, L.value


@pytest.mark.parametrize("build", [1, 2])
@pytest.mark.parametrize("hoist,leaf_max", [(0.01, 4), (-1.0, 10), (0.01, 1)])
@pytest.mark.parametrize("width", [2, 4])
@pytest.mark.parametrize("scene", ["cornell", "ico"])
def test_trace_matches_oracle(width, scene, hoist, leaf_max, build):
    """The GPU tree (the library's own, build 1, or the reference's re-laid
    out, build 2) against the oracle walking the reference's tree."""
    if scene == "cornell":
        s = Y.gen_cornell()
        lo, hi = (-0.99, -0.99, -3.99), (0.99, 0.99, -2.01)
    else:
        s = Y.gen_icosphere_room(11, 6, 3)
        lo, hi = (-4.9, -4.9, -4.9), (4.9, 4.9, 4.9)
    nodes = s.bvh()
    rays = random_rays(20000, 5, lo, hi)
    with Y.Context(bvh_width=width, max_paths=1 << 16, hoist_frac=hoist, leaf_max=leaf_max, bvh_build=build) as c:
        sid = c.scene(s, nodes)
        assert (c.scene_info(sid).n_hoist > 0) == (hoist > 0)
        gtri, glight, gt = c.trace(sid, rays)
    otri, olight, ot, oflags = oscene(s, nodes).trace(rays)
    shadow = rays[:, 7] != 0
    closest = ~shadow
    # closest hits
    same = (gtri == otri) & (glight == olight)
    bad = np.nonzero(closest & ~same)[0]
    ties = 0
    for i in bad:
        assert gtri[i] >= 0 and otri[i] >= 0, f"ray {i}: gpu {gtri[i]}/{glight[i]} oracle {otri[i]}/{olight[i]}"
        _, ta = tri_t(s, gtri[i], rays[i])
        _, tb = tri_t(s, otri[i], rays[i])
        assert abs(ta - tb) <= 1e-5 * max(ta, tb), f"ray {i}: not a tie ({ta} vs {tb})"
        ties += 1
    # every mismatch is a verified tie; ties come from triangles sharing an
    # edge or a plane and must stay a small minority
    assert ties <= 0.005 * closest.sum()
    ok = closest & same
    assert np.array_equal(gt[ok].view(np.int32), ot[ok].view(np.int32)), "t differs"
    # shadow rays: occlusion flag; occluder must really block
    gocc = (gtri >= 0) | (glight >= 0)
    oocc = (otri >= 0) | (olight >= 0)
    assert np.array_equal(gocc[shadow], oocc[shadow])
    for i in np.nonzero(shadow & (gtri >= 0))[0][:500]:
        got, t = tri_t(s, gtri[i], rays[i])
        assert got and t < rays[i, 3]


def test_brute_force_matches_bvh():
    s = Y.gen_cornell()
    rays = random_rays(4000, 9, (-0.99, -0.99, -3.99), (0.99, 0.99, -2.01), shadow_frac=0.0)
    with Y.Context(max_paths=1 << 14) as c:
        a = c.trace(c.scene(s), rays)
        b = c.trace(c.scene(s, brute=True), rays)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[2].view(np.int32), b[2].view(np.int32))


def _cornell_job(c, s, W, H, p, seed):
    sid = c.scene(s)
    return c.job(sid, s.cam, W, H, p, base_seed=seed)


@pytest.mark.parametrize("mis", [0, 1])
def test_primary_and_shadow_rays_match(mis):
    s = Y.gen_cornell()
    W, H = 96, 80
    p = Y.params_reference(mis=mis, max_bounces=8, rr_threshold=2)
    with Y.Context(max_paths=1 << 16) as c:
        job = _cornell_job(c, s, W, H, p, 3)
        cap = job.capture_next()
        job.render(1)
    seeds = O.mt19937(3, 1)
    _, ocap, _, _ = oscene(s, s.bvh(), heap=1500).render(ocam(s.cam), W, H, O.params(0, mis=mis, max_bounces=8,
                                                         rr_threshold=2), seeds, capture=True, threads=8)
    assert np.array_equal(cap["prim_tri"], ocap["prim_tri"])
    assert np.array_equal(cap["prim_light"], ocap["prim_light"])
    hit = cap["prim_tri"] >= 0
    assert np.array_equal(cap["prim_t"][hit].view(np.int32), ocap["prim_t"][hit].view(np.int32))
    rec = ocap["shadow_occ"] >= 0
    assert np.array_equal(cap["shadow_occ"] >= 0, rec)
    diff = cap["shadow_ray"][rec, :7].view(np.int32) != ocap["shadow_ray"][rec, :7].view(np.int32)
    assert not diff.any(), f"shadow rays differ: {diff.sum(0)} of {rec.sum()}"
    assert np.array_equal(cap["shadow_occ"][rec], ocap["shadow_occ"][rec])
    assert hit.mean() > 0.5 and rec.mean() > 0.4


def identical(a, b):
    return float((a.view(np.int32) == b.view(np.int32)).all(-1).mean())


def mre(a, b):
    a = a[..., :3].astype(np.float64)
    b = b[..., :3].astype(np.float64)
    fin = np.isfinite(a).all(-1) & np.isfinite(b).all(-1)
    num = np.abs(a - b).sum(-1)
    den = np.maximum(np.abs(b).sum(-1), 1e-2)
    return float((num[fin] / den[fin]).mean()), float(1 - fin.mean())


@pytest.mark.parametrize("mis,fix", [(0, 0), (1, 0), (1, 1)])
def test_radiance_cornell(mis, fix):
    s = Y.gen_cornell()
    W, H, F = 64, 64, 16
    kw = dict(mis=mis, fix_mis=fix, max_bounces=8, rr_threshold=2)
    with Y.Context(max_paths=1 << 16) as c:
        job = _cornell_job(c, s, W, H, Y.params_reference(**kw), 7)
        job.render(F)
        gimg = job.read()
    oimg, _, _, _ = oscene(s, s.bvh()).render(ocam(s.cam), W, H, O.params(0, **kw), O.mt19937(7, F), threads=8)
    assert np.array_equal(gimg[..., 3], oimg[..., 3])        # sample counts
    assert identical(gimg, oimg) >= 0.999
    e, nonfinite = mre(gimg, oimg)
    assert e <= 0.01, e
    assert nonfinite <= 0.01


def test_radiance_icosphere_room():
    s = Y.gen_icosphere_room(3, 9, 2)
    W, H, F = 80, 45, 8
    kw = dict(mis=1, max_bounces=8, rr_threshold=2)
    with Y.Context(max_paths=1 << 16) as c:
        sid = c.scene(s)
        job = c.job(sid, s.cam, W, H, Y.params_reference(**kw), base_seed=11)
        job.render(F)
        gimg = job.read()
    oimg, _, _, _ = oscene(s, s.bvh()).render(ocam(s.cam), W, H, O.params(0, **kw), O.mt19937(11, F), threads=8)
    assert identical(gimg, oimg) >= 0.999
    e, nonfinite = mre(gimg, oimg)
    assert e <= 0.01, e


def test_full_size_scene_matches_oracle_on_a_tile():
    """The headline configuration at its full size: 983,054 triangles,
    1920x1080, NEE+MIS, depth 8, RR from bounce 3.  Two frames on the GPU
    (its own SAH tree, default launch options) against the oracle walking
    the reference's tree as the reference does (breadth-first list capped
    at HEAP_SIZE = 1500, udpt.cl:8, 288-324) on the 96x96 centre tile of the
    same two frames: primary hits of the first frame and pixels.  No query
    of the tile reaches the cap, so the capped walk finds the exact closest
    hits the GPU's traversal finds."""
    s = Y.gen_icosphere_room(2024, 48, 5)
    nodes = s.bvh()
    W, H, F = 1920, 1080, 2
    kw = dict(mis=1, max_bounces=8, rr_threshold=2)
    with Y.Context(max_paths=1 << 23) as c:
        job = c.job(c.scene(s, nodes), s.cam, W, H, Y.params_reference(**kw), base_seed=5)
        cap = job.capture_next()
        job.render(F)
        gimg = job.read()
    x0, y0 = W // 2 - 48, H // 2 - 48
    rect = (x0, y0, x0 + 96, y0 + 96)
    osc = oscene(s, nodes, heap=1500)
    oimg, ocap, rays, _ = osc.render(ocam(s.cam), W, H, O.params(0, **kw), O.mt19937(5, F), rect=rect, threads=8,
                                     capture=True)
    assert rays > 96 * 96 * F and osc.last_counts["truncated"] == 0, osc.last_counts
    tile = np.zeros((H, W), bool)
    tile[y0:y0 + 96, x0:x0 + 96] = True
    t = tile.reshape(-1)
    same_tri = (cap["prim_tri"][t] == ocap["prim_tri"][t]).mean()
    assert same_tri >= 0.999, same_tri
    hit = t & (cap["prim_tri"] == ocap["prim_tri"]) & (cap["prim_tri"] >= 0)
    assert np.array_equal(cap["prim_t"][hit].view(np.int32), ocap["prim_t"][hit].view(np.int32))
    g, o = gimg[y0:y0 + 96, x0:x0 + 96], oimg[y0:y0 + 96, x0:x0 + 96]
    assert np.array_equal(g[..., 3], o[..., 3])        # sample counts
    assert identical(g, o) >= 0.999
    e, nonfinite = mre(g, o)
    assert e <= 0.01, e
    # the whole frame: every pixel has its samples and a finite colour
    assert (gimg[..., 3] == F).all()
    assert np.isfinite(gimg).all(-1).mean() >= 0.999


@pytest.mark.parametrize("sigma2", [0.0, 0.1, 0.5])
def test_radiance_oren_nayar(sigma2):
    """Oren-Nayar diffuse (udpt-primitives.cl:673-725, 1147-1210; used by
    evaluateBRDF for a pure diffuse lobe, rr_prob 1) with sigma^2 =
    alpha_x on the Cornell box's diffuse walls, against the oracle.  For
    sigma^2 > 0 the image differs from Lambert's, so the branch ran."""
    s = Y.gen_cornell()
    diffuse = (s.mats["ks"][:, :3].max(1) == 0) & (s.mats["is_specular"] == 0)
    s.mats["alpha_x"][diffuse] = sigma2
    W, H, F = 64, 64, 16
    kw = dict(mis=1, max_bounces=8, rr_threshold=2)
    with Y.Context(max_paths=1 << 16) as c:
        sid = c.scene(s)
        job = c.job(sid, s.cam, W, H, Y.params_reference(oren_nayar=1, **kw), base_seed=13)
        job.render(F)
        gimg = job.read()
        lam = c.job(sid, s.cam, W, H, Y.params_reference(**kw), base_seed=13)
        lam.render(F)
        limg = lam.read()
    oimg, _, _, _ = oscene(s, s.bvh()).render(ocam(s.cam), W, H, O.params(0, oren_nayar=1, **kw), O.mt19937(13, F),
                                              threads=8)
    assert np.array_equal(gimg[..., 3], oimg[..., 3])
    assert identical(gimg, oimg) >= 0.999
    e, nonfinite = mre(gimg, oimg)
    assert e <= 0.01 and nonfinite <= 0.001, (e, nonfinite)
    if sigma2 > 0:
        assert mre(gimg, limg)[0] > 0.01


def reference_cornellbox():
    """geometry/cornellbox.obj + .mtl of the reference (diffuse walls with
    newMaterial's alpha_x = 100, a glossy box, a mirror and a
    mirror-transmissive material), as the arrays Scene::loadModel makes:
    tests/golden/cornellbox_scene.npz."""
    d = np.load(os.path.join(GOLDEN, "cornellbox_scene.npz"))
    return Y.Scene(d["tris"].copy(), d["mats"].copy(), Y.LightQuad.from_buffer_copy(d["light"].tobytes()),
                   Y.Cam.from_buffer_copy(d["cam"].tobytes()))


# geometry/cornellbox.obj uses only the grey, red and green materials of its
# .mtl; the `boxes` variants give its two boxes (triangles 0-9: shortBox,
# 10-19: longBox, in file order) other materials of the same .mtl so that the
# glossy, mirror and mirror-transmissive entries are rendered too.
BOX_MATS = {"glossy/mirror-transmissive": (6, 4), "teapot/mirror": (3, 5)}


@pytest.mark.parametrize("mode,kw,boxes", [
    (Y.YN_MODE_PT, dict(), None),
    (Y.YN_MODE_PT, dict(mis=1, max_bounces=8, rr_threshold=2), None),
    (Y.YN_MODE_PT, dict(mis=1, max_bounces=8, rr_threshold=2), "glossy/mirror-transmissive"),
    (Y.YN_MODE_PT, dict(mis=1, max_bounces=8, oren_nayar=1), "teapot/mirror"),
    (Y.YN_MODE_BDPT, dict(mis=1, bdpt_bounces=5), "glossy/mirror-transmissive")])
def test_reference_cornellbox_obj_matches_oracle(mode, kw, boxes):
    """The reference's own scene file, rendered on the GPU and by the
    oracle walking the reference's BVH of it (the golden node array)."""
    s = reference_cornellbox()
    if boxes is not None:
        s.tris["matID"][0:10], s.tris["matID"][10:20] = BOX_MATS[boxes]
    nodes = np.load(os.path.join(GOLDEN, "cornellbox_nodes.npz"))["nodes"]
    W, H, F = 80, 60, 8
    with Y.Context(max_paths=1 << 16) as c:
        job = c.job(c.scene(s, nodes), s.cam, W, H, Y.params_reference(mode, **kw), base_seed=17)
        cap = job.capture_next()
        job.render(F)
        gimg = job.read()
    oimg, ocap, _, _ = oscene(s, nodes, heap=1500).render(ocam(s.cam), W, H, O.params(mode, **kw),
                                                          O.mt19937(17, F), threads=8, capture=True)
    assert np.array_equal(cap["prim_tri"], ocap["prim_tri"])
    assert np.array_equal(gimg[..., 3], oimg[..., 3])
    assert identical(gimg, oimg) >= 0.999
    e, nonfinite = mre(gimg, oimg)
    assert e <= 0.01 and nonfinite <= 0.001, (e, nonfinite)
    if boxes is not None:
        # both boxes' materials are seen by primary rays
        seen = set(s.tris["matID"][cap["prim_tri"][cap["prim_tri"] >= 0]].tolist())
        assert set(BOX_MATS[boxes]) <= seen, seen


def test_reference_defaults_unbounded_loop():
    """udpt.cl as shipped: no bounce limit, RR after bounce 6, MIS off."""
    s = Y.gen_cornell()
    W, H, F = 32, 32, 4
    with Y.Context(max_paths=1 << 14) as c:
        job = _cornell_job(c, s, W, H, Y.params_reference(), 5)
        job.render(F)
        gimg = job.read()
    oimg, _, _, _ = oscene(s, s.bvh()).render(ocam(s.cam), W, H, O.params(0), O.mt19937(5, F), threads=8)
    assert identical(gimg, oimg) >= 0.999
    e, _ = mre(gimg, oimg)
    assert e <= 0.01, e


def comb_scene(n):
    """n small triangles along the -z axis, under a hand-made comb-shaped
    BVH (every interior node: one leaf child and one interior child), so
    a ray down the axis keeps one deferred child per level on its stack:
    deeper than the LDS part of the traversal stack."""
    s = Y.gen_cornell()
    tris = np.zeros(n, dtype=Y.TRI_DTYPE)
    for i in range(n):
        z = -1.0 - 0.1 * i
        tris[i]["v1"] = (-0.05, -0.05, z, 1)
        tris[i]["v2"] = (0.05, -0.05, z, 1)
        tris[i]["v3"] = (0.0, 0.05, z, 1)
        for k in ("vn1", "vn2", "vn3"):
            tris[i][k] = (0, 0, 1, 0)
        tris[i]["matID"] = 0
    def tri_box(i):
        v = np.stack([tris[i]["v1"][:3], tris[i]["v2"][:3], tris[i]["v3"][:3]])
        lo, hi = v.min(0), v.max(0)
        hi = np.where(hi - lo == 0, hi + np.float32(0.2), hi)
        return lo, hi
    # interior node j (index 2j, j < n-1): children 2j+1, a leaf holding the
    # far triangle n-1-j, and 2j+2, the next interior node (the last index
    # is the leaf of the nearest triangle 0)
    out = np.zeros(2 * n - 1, dtype=Y.NODE_DTYPE)
    out["vert_list"] = -1
    for j in range(n - 1):
        out[2 * j]["child_idx"] = 2 * j + 1
        out[2 * j]["vert_len"] = -1
        out[2 * j + 1]["child_idx"] = -1
        out[2 * j + 1]["vert_len"] = 1
        out[2 * j + 1]["vert_list"][0] = n - 1 - j
    out[2 * (n - 1)]["child_idx"] = -1
    out[2 * (n - 1)]["vert_len"] = 1
    out[2 * (n - 1)]["vert_list"][0] = 0
    for j in range(2 * n - 2, -1, -1):
        nd = out[j]
        if nd["vert_len"] > 0:
            lo, hi = tri_box(nd["vert_list"][0])
        else:
            c = nd["child_idx"]
            lo = np.minimum(out[c]["p_min"][:3], out[c + 1]["p_min"][:3])
            hi = np.maximum(out[c]["p_max"][:3], out[c + 1]["p_max"][:3])
        out[j]["p_min"] = (*lo, 1)
        out[j]["p_max"] = (*hi, 1)
    s.tris, s.nodes = tris, out
    s.mats = s.mats[:1]
    return s


@pytest.mark.parametrize("width", [2, 4])
def test_deep_tree_uses_stack_overflow_area(width):
    s = comb_scene(60)
    rays = np.zeros((64, 8), np.float32)
    rays[:, 2] = 0.0
    rays[:, 0] = np.linspace(-0.01, 0.01, 64)
    rays[:, 3] = np.inf
    rays[:, 6] = -1.0
    with Y.Context(bvh_width=width, max_paths=1 << 12, hoist_frac=-1, leaf_max=10, bvh_build=2) as c:
        sid = c.scene(s, s.nodes)
        assert c.scene_info(sid).max_depth > 16
        gtri, glight, gt = c.trace(sid, rays)
    otri, olight, ot, _ = oscene(s, s.nodes).trace(rays)
    assert np.array_equal(gtri, otri) and np.array_equal(gt.view(np.int32), ot.view(np.int32))
    assert (gtri == 0).all()


@pytest.mark.parametrize("graph", [0, 1])
def test_regeneration_with_work_beyond_capacity(graph):
    """16 frames of 64x64 through 4096 paths in flight: new paths are
    appended behind the queued rays as room frees up, the bounce loop also
    replayed from a graph, and the image is still the oracle's bit for bit."""
    s = Y.gen_cornell()
    W, H, F = 64, 64, 16
    kw = dict(mis=1, max_bounces=8, rr_threshold=2)
    with Y.Context(max_paths=1 << 12, batch_factor=2, graph=graph) as c:
        job = _cornell_job(c, s, W, H, Y.params_reference(**kw), 21)
        job.render(F)
        gimg = job.read()
        st = job.stats()
    oimg, _, _, _ = oscene(s, s.bvh()).render(ocam(s.cam), W, H, O.params(0, **kw), O.mt19937(21, F), threads=8)
    assert st.samples == W * H * F
    assert np.array_equal(gimg[..., 3], oimg[..., 3])
    assert identical(gimg, oimg) >= 0.999


@pytest.mark.parametrize("mode", [Y.YN_MODE_PT, Y.YN_MODE_BDPT])
def test_lds_stack_sizes_keep_images(mode):
    """The traversal stack's LDS part (8, 12, 16 or 32 entries, or chosen
    per launch) only moves entries between LDS and the overflow area in
    memory: images and node counts are the same bit for bit."""
    s = Y.gen_icosphere_room(3, 9, 2)
    W, H, F = 64, 48, 2
    kw = dict(bdpt_bounces=4) if mode else dict(mis=1, max_bounces=8, rr_threshold=2)
    p = Y.params_reference(mode, **kw)
    out = []
    for lds in (0, 8, 12, 16, 32):
        with Y.Context(max_paths=1 << 14, lds_stack=lds) as c:
            assert c.config()["lds_stack"] == lds
            job = c.job(c.scene(s), s.cam, W, H, p, base_seed=29)
            job.set_counting(True)
            job.render(F)
            st = job.stats()
            out.append((job.read(), st.nodes_ext, st.nodes_shadow))
    for img, ne, ns in out[1:]:
        assert np.array_equal(img.view(np.int32), out[0][0].view(np.int32))
        assert (ne, ns) == out[0][1:]


@pytest.mark.parametrize("mode", [Y.YN_MODE_PT, Y.YN_MODE_BDPT])
def test_trace_overlap_keeps_images(mode):
    """Closest-hit and any-hit traces on two streams (each with its own half
    of the stack-overflow area, joined before the shade) give the same image
    as on one stream, bit for bit."""
    s = Y.gen_icosphere_room(3, 9, 2)
    W, H, F = 64, 48, 3
    kw = dict(bdpt_bounces=4) if mode else dict(mis=1, max_bounces=8, rr_threshold=2)
    p = Y.params_reference(mode, **kw)
    imgs = []
    for ov in (0, 1):
        with Y.Context(max_paths=1 << 14, trace_overlap=ov, lds_stack=8) as c:
            assert c.config()["trace_overlap"] == ov
            job = c.job(c.scene(s), s.cam, W, H, p, base_seed=17)
            job.render(F)
            imgs.append(job.read())
    assert np.array_equal(imgs[0].view(np.int32), imgs[1].view(np.int32))


@pytest.mark.parametrize("mode", [Y.YN_MODE_PT, Y.YN_MODE_BDPT])
def test_graph_replay_keeps_images(mode):
    """The bounce loop replayed from a captured HIP graph gives the same
    images as launched kernel by kernel, bit for bit: over many replays of one
    graph (frames rendered one call at a time, with host work and
    synchronisation between them) and when the arguments change (two scenes
    and sizes alternating, which captures again)."""
    scenes = [Y.gen_icosphere_room(3, 9, 2), Y.gen_cornell()]
    sizes = [(64, 48), (40, 40)]
    kw = dict(bdpt_bounces=4) if mode else dict(mis=1, max_bounces=8, rr_threshold=2)
    p = Y.params_reference(mode, **kw)
    imgs = []
    for gr in (0, 1):
        with Y.Context(max_paths=1 << 14, graph=gr, lds_stack=8) as c:
            assert c.config()["graph"] == gr
            jobs = [c.job(c.scene(s), s.cam, W, H, p, base_seed=23 + k)
                    for k, (s, (W, H)) in enumerate(zip(scenes, sizes))]
            for _ in range(3):
                for j in jobs:
                    j.render(1)
            jobs[0].render(2)
            imgs.append([j.read() for j in jobs])
    for a, b in zip(*imgs):
        assert np.array_equal(a.view(np.int32), b.view(np.int32))


def test_nan_rays_hit_nothing_like_the_oracle():
    """Rays with a NaN in origin, direction or length: every triangle test
    rejects them (NaN u, v or t), so both kinds of query answer "no hit",
    as the oracle's traversal does; the trace answers them without walking
    the tree (the reference's NaN guard would keep every box)."""
    s = Y.gen_icosphere_room(11, 6, 3)
    rays = random_rays(512, 3, (-4.9, -4.9, -4.9), (4.9, 4.9, 4.9))
    for j, i in enumerate(range(0, 512, 8)):
        rays[i, j % 7] = np.nan          # o.x, o.y, o.z, length, d.x, d.y, d.z in turn
    nan_rows = np.isnan(rays[:, :7]).any(1)
    assert nan_rows.sum() >= 60
    with Y.Context(max_paths=1 << 14) as c:
        t0 = time.perf_counter()
        gtri, glight, gt = c.trace(c.scene(s), rays)
        assert time.perf_counter() - t0 < 10
    otri, olight, ot, _ = oscene(s, s.bvh()).trace(rays)
    assert (gtri[nan_rows] == -1).all() and (otri[nan_rows] == -1).all()
    assert np.array_equal(glight[nan_rows], olight[nan_rows])
    ok = ~nan_rows
    assert (gtri[ok] == otri[ok]).mean() > 0.99


@pytest.mark.parametrize("mode", [Y.YN_MODE_PT, Y.YN_MODE_BDPT])
def test_trace_queue_slices_keep_images(mode):
    """Ray queues cut into slices (one work counter per XCD, waves moving on
    to the next slices once theirs is used up) trace the same rays: images
    are bit-identical for 1, 3 and 8 slices, and so are the trace results of
    caller rays."""
    s = Y.gen_icosphere_room(3, 9, 2)
    W, H, F = 64, 48, 3
    kw = dict(bdpt_bounces=4) if mode else dict(mis=1, max_bounces=8, rr_threshold=2)
    p = Y.params_reference(mode, **kw)
    rays = random_rays(12000, 8, (-4.9, -4.9, -4.9), (4.9, 4.9, 4.9))     # within the 16k paths
    imgs, traced = [], []
    for ns in (1, 3, 8):
        with Y.Context(max_paths=1 << 14, trace_xcd=ns) as c:
            assert c.config()["trace_xcd"] == ns
            sid = c.scene(s)
            job = c.job(sid, s.cam, W, H, p, base_seed=29)
            job.render(F)
            imgs.append(job.read())
            traced.append(c.trace(sid, rays))
    for k in (1, 2):
        assert np.array_equal(imgs[0].view(np.int32), imgs[k].view(np.int32))
        for a, b in zip(traced[0], traced[k]):
            assert np.array_equal(np.asarray(a).view(np.int32), np.asarray(b).view(np.int32))


@pytest.mark.parametrize("mode", [Y.YN_MODE_PT, Y.YN_MODE_BDPT])
def test_trace_refill_keeps_images_and_counts(mode):
    """Trace waves whose finished lanes take new rays (from one idle lane to
    a whole wave) give the images of waves that run their rays to the end,
    bit for bit, with two scenes in one batch (per-lane scene lookup), and
    visit exactly the same nodes and triangles (counting mode); caller rays
    (hoisted slots, brute force) trace the same too."""
    scenes = [Y.gen_icosphere_room(3, 9, 2), Y.gen_cornell()]
    kw = dict(bdpt_bounces=4) if mode else dict(mis=1, max_bounces=8, rr_threshold=2)
    p = Y.params_reference(mode, **kw)
    rays = random_rays(12000, 4, (-4.9, -4.9, -4.9), (4.9, 4.9, 4.9))
    res = {}
    for rf in (-1, 1, 16, 64):
        with Y.Context(max_paths=1 << 14, trace_refill=rf) as c:
            assert c.config()["trace_refill"] == rf
            sids = [c.scene(s) for s in scenes]
            jobs = [c.submit(sids[k], s.cam, 48, 40, p, 2, base_seed=31 + k) for k, s in enumerate(scenes)]
            for _ in range(100):
                if c.serve(4) == 0:
                    break
            imgs = [j.read() for j in jobs]
            cj = c.job(sids[0], scenes[0].cam, 40, 32, p, base_seed=5)
            cj.set_counting(True)
            cj.render(1)
            st = cj.stats()
            tr = c.trace(sids[0], rays) + c.trace(c.scene(scenes[1], brute=True), rays[:4000])
            res[rf] = (imgs, (st.rays_ext, st.nodes_ext, st.tris_ext, st.rays_shadow, st.nodes_shadow,
                              st.tris_shadow), tr)
    for rf in (1, 16, 64):
        for a, b in zip(res[-1][0], res[rf][0]):
            assert np.array_equal(a.view(np.int32), b.view(np.int32)), rf
        assert res[rf][1] == res[-1][1], rf
        for a, b in zip(res[-1][2], res[rf][2]):
            assert np.array_equal(np.asarray(a).view(np.int32), np.asarray(b).view(np.int32)), rf
    assert res[-1][1][1] > 0 and res[-1][1][4] > 0


def test_fma_slab_form_keeps_full_size_images(monkeypatch):
    """The traversal's FMA slab form against the reference's slab arithmetic
    on the headline scene (983,054 triangles): the same job rendered on one
    upload of the scene whose rays use the FMA form and one that keeps every
    ray on the reference's form (YUNE_SLAB_REFERENCE=1) gives the same
    primary hits and the same image bit for bit, with node visits equal to
    within 0.1%."""
    s = Y.gen_icosphere_room(2024, 48, 5)
    nodes = s.bvh()
    W, H, F = 640, 360, 2
    p = Y.params_reference(mis=1, max_bounces=8, rr_threshold=2)
    out = {}
    with Y.Context(max_paths=1 << 22) as c:
        for form in ("reference", "fma"):
            if form == "reference":
                monkeypatch.setenv("YUNE_SLAB_REFERENCE", "1")
            else:
                monkeypatch.delenv("YUNE_SLAB_REFERENCE", raising=False)
            job = c.job(c.scene(s, nodes), s.cam, W, H, p, base_seed=9)
            job.set_counting(True)
            cap = job.capture_next()
            job.render(F)
            out[form] = (cap, job.read(), job.stats())
    (ca, ia, sa), (cb, ib, sb) = out["reference"], out["fma"]
    print(f"  nodes per ray: reference {sa.nodes_ext / max(sa.rays_ext, 1):.3f}, "
          f"fma {sb.nodes_ext / max(sb.rays_ext, 1):.3f}", flush=True)
    assert np.array_equal(ca["prim_tri"], cb["prim_tri"])
    assert np.array_equal(ca["prim_t"].view(np.int32), cb["prim_t"].view(np.int32))
    assert np.array_equal(ia.view(np.int32), ib.view(np.int32))
    assert sa.rays_ext == sb.rays_ext and sa.rays_shadow == sb.rays_shadow
    # the same nodes, up to the few boxes the widened test also enters and
    # visit-order changes among near-equal entry distances
    assert abs(int(sb.nodes_ext) - int(sa.nodes_ext)) <= 0.001 * sa.nodes_ext
