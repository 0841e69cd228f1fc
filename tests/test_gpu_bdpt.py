"""GPU parity of the HIP bidirectional tracer (bdpt.cl semantics) against
the CPU oracle, plus a closed form.

Tolerances:
  * camera-ray hits: same triangle / light id and t, bit for bit;
  * the first NEE shadow ray (eye vertex 0): recorded for the same pixels
    and bit-identical;
  * radiance: bit-identical in at least 99.9% of pixels, and mean over
    pixels of |gpu - cpu|_1 / max(|cpu|_1, 1e-2) <= 1%.  Kernels and oracle
    share one definition of sin, cos and pow (yn_fmath.h) and every other
    float operation rounds once in the reference's order on both sides, so
    a path's random decisions and arithmetic agree exactly; only a closest
    hit at equal t on two triangles (resolved by test order) may differ.
"""
import numpy as np
import pytest

import yune_amd as Y
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def oscene(s, nodes, heap=0):
    return O.OScene(s.tris, s.mats, nodes, s.light, heap_size=heap)


def ocam(cam):
    return O.Cam.from_buffer_copy(bytes(cam))


def identical(a, b):
    return float((a.view(np.int32) == b.view(np.int32)).all(-1).mean())


def mre(a, b):
    a = a[..., :3].astype(np.float64)
    b = b[..., :3].astype(np.float64)
    fin = np.isfinite(a).all(-1) & np.isfinite(b).all(-1)
    num = np.abs(a - b).sum(-1)
    den = np.maximum(np.abs(b).sum(-1), 1e-2)
    return float((num[fin] / den[fin]).mean()), float(1 - fin.mean())


def bdpt_params(**kw):
    kw.setdefault("bdpt_bounces", 5)
    return kw


def test_bdpt_camera_hits_and_first_shadow_ray():
    s = Y.gen_cornell()
    W, H = 96, 80
    kw = bdpt_params(mis=1)
    with Y.Context(max_paths=1 << 16) as c:
        job = c.job(c.scene(s), s.cam, W, H, Y.params_reference(Y.YN_MODE_BDPT, **kw), base_seed=3)
        cap = job.capture_next()
        job.render(1)
    _, ocap, _, _ = oscene(s, s.bvh(), heap=1500).render(ocam(s.cam), W, H, O.params(1, **kw), O.mt19937(3, 1),
                                                         capture=True, threads=8)
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


@pytest.mark.parametrize("mis,fix", [(0, 0), (1, 0), (1, 1)])
def test_bdpt_radiance_cornell(mis, fix):
    s = Y.gen_cornell()
    W, H, F = 64, 64, 16
    kw = bdpt_params(mis=mis, fix_mis=fix)
    with Y.Context(max_paths=1 << 16) as c:
        job = c.job(c.scene(s), s.cam, W, H, Y.params_reference(Y.YN_MODE_BDPT, **kw), base_seed=7)
        job.render(F)
        gimg = job.read()
    oimg, _, _, _ = oscene(s, s.bvh()).render(ocam(s.cam), W, H, O.params(1, **kw), O.mt19937(7, F), threads=8)
    assert np.array_equal(gimg[..., 3], oimg[..., 3])
    assert identical(gimg, oimg) >= 0.999
    e, nonfinite = mre(gimg, oimg)
    assert e <= 0.01, e
    assert nonfinite <= 0.01


def test_bdpt_reference_defaults_and_row_bands():
    """bdpt.cl as shipped (BDPT_BOUNCES 20, RR after vertex 4, MIS off),
    with a path capacity below one frame so the image runs in row bands."""
    s = Y.gen_cornell()
    W, H, F = 48, 40, 4
    with Y.Context(max_paths=700) as c:
        job = c.job(c.scene(s), s.cam, W, H, Y.params_reference(Y.YN_MODE_BDPT), base_seed=5)
        job.render(F)
        gimg = job.read()
    oimg, _, _, _ = oscene(s, s.bvh()).render(ocam(s.cam), W, H, O.params(1), O.mt19937(5, F), threads=8)
    assert identical(gimg, oimg) >= 0.999
    e, _ = mre(gimg, oimg)
    assert e <= 0.01, e


def box_scene(a, e, light_ke=(5, 5, 5)):
    """Closed grey cube [-1,1]^3 (normals inward, albedo a, emission e);
    the light quad emits but sits outside, facing away: neither its light
    subpaths nor NEE from inside reach the box."""
    corners = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], float)
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = np.zeros(12, Y.TRI_DTYPE)
    k = 0
    for f in faces:
        q = corners[list(f)]
        n = -q.mean(0) / np.linalg.norm(q.mean(0))
        for i0, i1, i2 in ((0, 1, 2), (0, 2, 3)):
            tris[k]["v1"], tris[k]["v2"], tris[k]["v3"] = (*q[i0], 1), (*q[i1], 1), (*q[i2], 1)
            tris[k]["vn1"] = tris[k]["vn2"] = tris[k]["vn3"] = (*n, 0)
            k += 1
    mats = np.zeros(1, Y.MAT_DTYPE)
    mats[0]["kd"] = (a, a, a, 1)
    mats[0]["ke"] = (e, e, e, 1)
    mats[0]["ks"] = (0, 0, 0, 1)
    mats[0]["n"] = mats[0]["k"] = mats[0]["px"] = mats[0]["py"] = 1
    light = Y.light_quad((5, 5, 5), (0, 1, 0), light_ke, (0.1, 0, 0), (0, 0, 0.1))
    cam = Y.camera((1, 0, 0), (0, 1, 0), (0, 0, -1), (0.1, 0.2, 0.3), 60.0)
    return Y.Scene(tris, mats, light, cam)


def closed_box_with_light():
    """The grey cube of box_scene with the light quad inside it, under the
    top face, facing down: light and eye subpaths never leave the box."""
    s = box_scene(0.5, 0.0)
    s.light = Y.light_quad((-0.2, 0.9, -0.2), (0, -1, 0), (5, 4, 3), (0.4, 0, 0), (0, 0, 0.4))
    return s


def test_bdpt_forty_vertex_subpaths():
    """bdpt_bounces = 40 with roulette off in a closed box: light subpaths
    of 40 vertices, so each eye vertex issues connection rays to light
    vertices j >= 32 (64-bit connection masks), against the oracle."""
    s = closed_box_with_light()
    W, H, F = 24, 24, 2
    kw = dict(mis=1, bdpt_bounces=40, rr_threshold=100)
    with Y.Context(max_paths=1 << 14) as c:
        job = c.job(c.scene(s), s.cam, W, H, Y.params_reference(Y.YN_MODE_BDPT, **kw), base_seed=19)
        job.render(F)
        gimg = job.read()
        st = job.stats()
    osc = oscene(s, s.bvh())
    oimg, _, _, _ = osc.render(ocam(s.cam), W, H, O.params(1, **kw), O.mt19937(19, F), threads=8)
    # 40 eye vertices per sample, each with connection rays to the ~22 of
    # its 39 light vertices that face it
    assert osc.last_counts["conn_traced"] > W * H * F * 40 * 15
    assert st.rays_shadow > W * H * F * 40 * 15
    assert identical(gimg, oimg) >= 0.999
    e, nonfinite = mre(gimg, oimg)
    assert e <= 0.01, e


def test_bdpt_furnace_closed_form():
    """Eye-subpath weights and emission of bdpt.cl shading, independent of
    the oracle: 2e (1 - (0.1 a)^B) / (1 - 0.1 a) in every pixel."""
    a, e, B = 0.5, 0.2, 5
    s = box_scene(a, e)
    with Y.Context(max_paths=1 << 14) as c:
        job = c.job(c.scene(s), s.cam, 24, 16, Y.params_reference(Y.YN_MODE_BDPT, bdpt_bounces=B, rr_threshold=100),
                    base_seed=2)
        job.render(2)
        img = job.read()
    exp = 2 * e * (1 - (0.1 * a) ** B) / (1 - 0.1 * a)
    assert np.allclose(img[..., :3], exp, rtol=2e-5)


def test_pt_furnace_closed_form():
    """The unidirectional counterpart: e (1 + (1 + a)(1 - a^B) / (1 - a))."""
    a, e, B = 0.5, 0.2, 3
    s = box_scene(a, e)
    with Y.Context(max_paths=1 << 14) as c:
        job = c.job(c.scene(s), s.cam, 24, 16, Y.params_reference(max_bounces=B, rr_threshold=100), base_seed=2)
        job.render(2)
        img = job.read()
    exp = e * (1 + (1 + a) * (1 - a ** B) / (1 - a))
    assert np.allclose(img[..., :3], exp, rtol=2e-5)


def test_device_fmath_is_bit_identical_to_host():
    """The kernels' sin, cos and pow on the GPU equal the host evaluation
    (the oracle's) bit for bit; this is what lets bounce directions, Phong
    lobes and Fresnel terms match the oracle exactly."""
    from yune_amd import _lib
    rng = np.random.default_rng(12)
    n = 200000
    x = np.concatenate([rng.uniform(0, 6.3, n), rng.uniform(-100, 100, n // 4)]).astype(np.float32)
    px = rng.uniform(0, 2, len(x)).astype(np.float32)
    py = np.concatenate([rng.uniform(0, 1000, len(x) // 2), rng.uniform(-4, 4, len(x) - len(x) // 2)]).astype(np.float32)
    L = O.lib()
    with Y.Context(max_paths=1 << 12) as c:
        for op, fn, a, b in ((0, L.or_fm_sinf, x, None), (1, L.or_fm_cosf, x, None), (2, L.or_fm_powf, px, py)):
            out = np.empty_like(a)
            c._ck(_lib.lib().yn_fmath_eval(c._p, op, _lib.ptr(a), _lib.ptr(b) if b is not None else None, len(a),
                                           _lib.ptr(out)))
            ref = np.array([fn(*v) for v in (zip(a, b) if b is not None else zip(a))], np.float32)
            assert np.array_equal(out.view(np.int32), ref.view(np.int32)), op
