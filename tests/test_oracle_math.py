"""Pins the oracle's shading and tracing pieces against mathematics.

Each check is chosen so that a plausible slip in the restatement (a dropped
cosine or distance term, a wrong normaliser, a flipped normal, a swapped
operand) fails: closed forms, hemisphere integrals by quadrature, sampling
moments, Snell's law, brute force over every triangle, and an emissive
"furnace" whose radiance is a geometric series.  Reference locations:
kernels/legacy/udpt.cl (ray tests :326-431, BRDFs :611-630, sampling
:700-910, Fresnel :912-1031, calcPhongPDF :1033-1045, sampleLights
:632-667, createRay :213-238), udpt-primitives.cl:696-725 (Oren-Nayar),
kernels/tonemap.cl:14-47, src/Camera.cpp:49.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate

import yune_amd as Y
from oracle import pyoracle as O

L = O.lib()
f4p = lambda a: np.ascontiguousarray(a, np.float32).ctypes.data_as(C.c_void_p)  # noqa: E731


def v4(x, y, z, w=0.0):
    return np.array([x, y, z, w], np.float32)


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def mat(**kw):
    m = np.zeros((), O.MAT_DTYPE)
    m["ke"] = (0, 0, 0, 1)
    m["kd"] = (0.3, 0.3, 0.3, 1)
    m["ks"] = (0, 0, 0, 1)
    m["n"] = m["k"] = m["px"] = m["py"] = 1
    m["alpha_x"] = m["alpha_y"] = 100
    for k, v in kw.items():
        m[k] = v
    return m


def mptr(m):
    m = np.ascontiguousarray(m)
    return m, m.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- rays
def ray_box(o, d, lo, hi):
    bb = np.array([*lo, 1, *hi, 1], np.float32)
    return L.or_ray_box(f4p(v4(*o, 1)), f4p(v4(*d)), f4p(bb))


def test_ray_box_hand_cases():
    lo, hi = (-1, -1, -1), (1, 1, 1)
    assert ray_box((0, 0, -5), (0, 0, 1), lo, hi)           # straight through
    assert not ray_box((0, 0, 5), (0, 0, 1), lo, hi)        # box behind the origin
    assert ray_box((0, 0, 0), (0.3, 0.2, 1), lo, hi)        # origin inside
    assert not ray_box((0, 2, -5), (0, 0, 1), lo, hi)       # parallel, outside the y slab
    assert ray_box((0, 1, -5), (0, 0, 1), lo, hi) in (0, 1)  # on the slab plane: 0*inf = NaN, axis skipped
    assert ray_box((0, 0.5, -5), (0, 0, 1), lo, hi)         # parallel inside the slab (+-inf bounds)
    assert not ray_box((-5, 3, 0), (1, 0.1, 0), lo, hi)     # passes above


def test_ray_box_against_exact_interval():
    """Random rays, compared with the slab interval in float64; cases within
    1e-4 of grazing are skipped (either answer is fine there)."""
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(4000):
        lo = rng.uniform(-2, 0, 3)
        hi = lo + rng.uniform(0.1, 2, 3)
        o = rng.uniform(-4, 4, 3)
        d = unit(rng.normal(size=3))
        with np.errstate(divide="ignore", invalid="ignore"):
            t1, t2 = (lo - o) / d, (hi - o) / d
        tn, tf = np.max(np.minimum(t1, t2)), np.min(np.maximum(t1, t2))
        if abs(tf - max(tn, 0)) < 1e-4 or abs(tf) < 1e-4:
            continue
        assert ray_box(o.astype(np.float32), d.astype(np.float32), lo.astype(np.float32),
                       hi.astype(np.float32)) == int(tf > max(tn, 0))
        checked += 1
    assert checked > 3500


def tri(a, b, c, n=(0, 0, 1)):
    t = np.zeros((), O.TRI_DTYPE)
    t["v1"], t["v2"], t["v3"] = (*a, 1), (*b, 1), (*c, 1)
    t["vn1"] = t["vn2"] = t["vn3"] = (*n, 0)
    return t


def ray_tri(t, o, d, length=np.inf):
    t = np.ascontiguousarray(t)
    ln = C.c_float(length)
    got = L.or_ray_tri(t.ctypes.data_as(C.c_void_p), f4p(v4(*o, 1)), f4p(v4(*d)), C.byref(ln))
    return got, ln.value


def test_ray_tri_against_linear_solve():
    """o + t d = a + u (b-a) + v (c-a), solved in float64 by numpy: hit iff
    u, v >= 0, u + v <= 1, 0 < t < len (no back-face culling)."""
    rng = np.random.default_rng(4)
    hits = 0
    for _ in range(3000):
        a, b, c = rng.uniform(-1, 1, (3, 3))
        o = rng.uniform(-3, 3, 3)
        uu, vv = rng.uniform(-0.3, 1.0, 2)               # aim near the triangle: about half hit
        d = unit(a + uu * (b - a) + vv * (c - a) - o) * (1 if rng.uniform() < 0.9 else -1)
        m = np.stack([d, a - b, a - c], 1)
        if abs(np.linalg.det(m)) < 1e-3:
            continue
        t, u, v = np.linalg.solve(m, a - o)
        margin = min(u, v, 1 - u - v, abs(t))
        if abs(margin) < 1e-4:
            continue
        exp = u >= 0 and v >= 0 and u + v <= 1 and t > 0
        got, tt = ray_tri(tri(a, b, c), o, d)
        assert got == exp
        if exp:
            hits += 1
            assert abs(tt - t) <= 1e-4 * max(1.0, abs(t))
            g2, t2 = ray_tri(tri(a, b, c), o, d, length=t * 0.99)   # farther than len: no hit, len kept
            assert not g2 and t2 == np.float32(t * 0.99)
    assert hits > 600


def test_ray_tri_back_face_and_edges():
    t = tri((0, 0, 0), (1, 0, 0), (0, 1, 0))
    assert ray_tri(t, (0.2, 0.2, 1), (0, 0, -1)) == (1, 1.0)
    assert ray_tri(t, (0.2, 0.2, -1), (0, 0, 1)) == (1, 1.0)    # from behind: still a hit
    assert ray_tri(t, (0.2, 0.2, 1), (0, 0, 1))[0] == 0         # pointing away
    assert ray_tri(t, (0.6, 0.6, 1), (0, 0, -1))[0] == 0        # outside u + v <= 1


# ---------------------------------------------------------------- camera
def test_camera_is_a_pinhole_with_the_stated_fov():
    """createRay on a look-at camera: pixel (x, y)'s ray crosses the image
    plane inside that pixel, the vertical half-angle is yfov/2 with pi read
    as 3.14 (Camera.cpp:49), and the eye is the origin."""
    side, up, look, eye = unit((1, 0, 1)), unit((0, 1, 0)), unit((-1, 0, 1)), np.array([0.5, 1.0, -2.0])
    cam = O.camera(side, up, look, eye, 45.0)
    W, H = 64, 48
    half = math.radians(45.0 / 2) * 3.14 / math.pi
    rng = np.random.default_rng(5)
    for _ in range(200):
        x, y, rand = int(rng.integers(W)), int(rng.integers(H)), int(rng.integers(2 ** 32))
        o, d = np.zeros(4, np.float32), np.zeros(4, np.float32)
        L.or_primary_ray(C.byref(cam), W, H, rand, x, y, f4p(o), f4p(d))
        o, d = o.copy(), d.copy()
        assert np.allclose(o[:3], eye, atol=1e-6) and abs(np.linalg.norm(d[:3]) - 1) < 1e-6
        fwd = d[:3] @ look
        assert fwd > 0
        u = (d[:3] @ side) / fwd / math.tan(half)         # in [-aspect, aspect]
        v = (d[:3] @ up) / fwd / math.tan(half)           # in [-1, 1], +1 = last row
        px, py = (u / (W / H) + 1) * W / 2, (v + 1) * H / 2
        seed = L.or_pixel_seed(rand, x, y, W)
        s1 = L.or_xor_shift(seed)
        s2 = L.or_xor_shift(s1)
        assert abs(px - (x + L.or_u2f(s1))) < 2e-3 and abs(py - (y + L.or_u2f(s2))) < 2e-3


# ---------------------------------------------------------------- sampling
def test_cosine_sampling_moments_and_pdf():
    n = unit((0.3, 0.9, -0.2)).astype(np.float32)
    nn = v4(*n)
    seed = C.c_uint32(777)
    d = np.zeros(4, np.float32)
    pdf = C.c_float()
    cos, pdfs = [], []
    for _ in range(20000):
        L.or_cosine_sample(f4p(v4(0, 0, 0, 1)), f4p(nn), C.byref(seed), f4p(d), C.byref(pdf))
        cos.append(float(d[:3] @ n))
        pdfs.append(pdf.value)
    cos, pdfs = np.array(cos), np.array(pdfs)
    assert (cos > -1e-6).all()
    assert np.allclose(pdfs, cos / math.pi, atol=2e-6)          # pdf = cos / pi
    assert abs(cos.mean() - 2 / 3) < 0.01                        # E[cos] under cos-weighted sampling
    assert abs((cos ** 2).mean() - 1 / 2) < 0.01


@pytest.mark.parametrize("px_py", [(5.0, 5.0), (3.2, 7.5)])
def test_phong_sampling_lobe(px_py):
    """Lobe around the mirror direction; cos^(e) with the exponent truncated
    to int(px + py); E[cos] = (e+1)/(e+2); pdf = (e+1)/(2pi) cos^e."""
    m, mp = mptr(mat(px=px_py[0], py=px_py[1], ks=(0.5, 0.5, 0.5, 1)))
    e = int(px_py[0] + px_py[1])
    n = v4(0, 0, 1)
    wi = unit((0.3, 0, 1)).astype(np.float32)
    refl = np.array([-wi[0], -wi[1], wi[2]])
    seed = C.c_uint32(99)
    d = np.zeros(4, np.float32)
    pdf = C.c_float()
    cs, ps = [], []
    for _ in range(20000):
        L.or_phong_sample(mp, f4p(v4(0, 0, 0, 1)), f4p(n), f4p(v4(*wi)), C.byref(seed), f4p(d), C.byref(pdf))
        if d[2] <= 0:
            assert pdf.value == 0
            continue
        cs.append(float(d[:3] @ refl))
        ps.append(pdf.value)
    cs, ps = np.array(cs), np.array(ps)
    assert abs(cs.mean() - (e + 1) / (e + 2)) < 0.01
    assert np.allclose(ps, (e + 1) / (2 * math.pi) * np.clip(cs, 0, 1) ** e, rtol=2e-3, atol=1e-5)


def brdf(m, n, wi, wo, glossy, rr=1.0, on=0):
    m, mp = mptr(m)
    out = np.zeros(4, np.float32)
    L.or_eval_brdf(mp, f4p(v4(*n)), f4p(v4(*wi)), f4p(v4(*wo)), glossy, C.c_float(rr), on, f4p(out))
    return out.copy()


def test_lambert_and_rr_prob():
    m = mat(kd=(0.6, 0.4, 0.2, 1))
    f = brdf(m, (0, 0, 1), unit((0.2, 0.1, 1)), unit((-0.4, 0.3, 1)), 0)
    assert np.allclose(f[:3], np.array([0.6, 0.4, 0.2]) / math.pi, rtol=1e-6)
    f2 = brdf(m, (0, 0, 1), unit((0.2, 0.1, 1)), unit((-0.4, 0.3, 1)), 0, rr=0.25)
    assert np.allclose(f2, f / 0.25, rtol=1e-6)


def hemi_integral(fn, n_t=200, n_p=400):
    """Integral of fn(w) cos(theta) over the upper hemisphere (midpoint rule)."""
    th = (np.arange(n_t) + 0.5) * (math.pi / 2) / n_t
    ph = (np.arange(n_p) + 0.5) * 2 * math.pi / n_p
    tot = 0.0
    for t in th:
        w = np.stack([np.sin(t) * np.cos(ph), np.sin(t) * np.sin(ph), np.full_like(ph, np.cos(t))], 1)
        vals = np.array([fn(x) for x in w])
        tot += vals.sum() * math.cos(t) * math.sin(t) * (math.pi / 2 / n_t) * (2 * math.pi / n_p)
    return tot


def test_glossy_brdf_normalisation():
    """At normal incidence the normalised Phong lobe (e+2)/(2pi) cos^e
    reflects exactly ks: integral f cos = ks (cos_alpha uses px+py, the
    normaliser int(px+py))."""
    m = mat(ks=(0.7, 0.7, 0.7, 1), px=10.0, py=10.0)
    n, wi = (0, 0, 1), (0, 0, 1)
    val = hemi_integral(lambda w: brdf(m, n, wi, w, 1)[0], n_t=120, n_p=64)
    assert abs(val - 0.7) < 0.005
    m2 = mat(ks=(1, 1, 1, 1), px=5.0, py=5.7)      # int(10.7) = 10 in the normaliser, 10.7 in the power
    w = unit((0.1, 0.2, 1))
    exp = (w[2] ** 10.7) * 12 / (2 * math.pi)
    assert abs(brdf(m2, n, wi, w, 1)[0] - exp) < 1e-5 * exp + 1e-7


def test_glossy_reflects_about_the_facing_normal():
    """w_i below the surface flips the normal before mirroring (udpt.cl reflect)."""
    m = mat(ks=(1, 1, 1, 1), px=20.0, py=20.0)
    wi = unit((0.3, 0, -1))
    mirror_flipped = unit((-0.3, 0, -1))
    assert brdf(m, (0, 0, 1), wi, mirror_flipped, 1)[0] > 1.0
    assert brdf(m, (0, 0, 1), wi, unit((-0.3, 0, 1)), 1)[0] < 1e-6


def oren_nayar_textbook(kd, sigma2, wi, wo):
    """Oren-Nayar qualitative model (Oren & Nayar 1994; PBRT 3e 8.4.2)."""
    ti, to = math.acos(wi[2]), math.acos(wo[2])
    pi_, po = math.atan2(wi[1], wi[0]), math.atan2(wo[1], wo[0])
    A = 1 - sigma2 / (2 * (sigma2 + 0.33))
    B = 0.45 * sigma2 / (sigma2 + 0.09)
    a, b = max(ti, to), min(ti, to)
    return kd / math.pi * (A + B * max(0.0, math.cos(pi_ - po)) * math.sin(a) * math.tan(b))


@pytest.mark.parametrize("sigma2", [0.0, 0.1, 0.5])
def test_oren_nayar_against_textbook(sigma2):
    m = mat(kd=(0.8, 0.8, 0.8, 1), alpha_x=sigma2)
    rng = np.random.default_rng(6)
    for _ in range(50):
        wi = unit(rng.normal(size=3) * (1, 1, 0) + (0, 0, rng.uniform(0.2, 2)))
        wo = unit(rng.normal(size=3) * (1, 1, 0) + (0, 0, rng.uniform(0.2, 2)))
        got = brdf(m, (0, 0, 1), wi, wo, 0, on=1)[0]
        assert abs(got - oren_nayar_textbook(0.8, sigma2, wi, wo)) < 2e-5
    # only with rr_prob == 1; other probabilities keep Lambert / rr_prob
    assert abs(brdf(m, (0, 0, 1), wi, wo, 0, rr=0.5, on=1)[0] - 0.8 / math.pi / 0.5) < 1e-6


def test_calc_phong_pdf_takes_cos_of_the_dot():
    """calcPhongPDF applies cos() to a cosine: at the mirror direction it is
    (e+1)/(2pi) cos(1)^e, not (e+1)/(2pi)."""
    m, mp = mptr(mat(px=2.0, py=2.0))
    n = v4(0, 0, 1)
    wo = v4(*unit((0.4, 0, 1)))
    mirror = v4(-wo[0], -wo[1], wo[2])
    got = L.or_phong_pdf(mp, f4p(n), f4p(mirror), f4p(wo))
    assert abs(got - 5 / (2 * math.pi) * math.cos(1.0) ** 4) < 1e-6


def fresnel(m, n, wi):
    m, mp = mptr(m)
    ior = C.c_float(-7)
    f = L.or_fresnel(mp, f4p(v4(*n)), f4p(v4(*wi)), C.byref(ior))
    return f, ior.value


def test_fresnel_cases():
    m = mat(n=1.5, ks=(0.04, 0.05, 0.06, 1))
    r0 = np.float32(np.float32(0.04) + np.float32(0.05) + np.float32(0.06)) / 3
    # entering at normal incidence: F = r0, ior factor n2^2/n1^2
    f, ior = fresnel(m, (0, 0, 1), (0, 0, 1))
    assert abs(f - r0) < 1e-7 and abs(ior - 2.25) < 1e-6
    # entering at 60 degrees: uses cos(theta_t) from Snell, F = r0 + (1-r0)(1 - cos_t^5)
    wi = (math.sin(math.pi / 3), 0, math.cos(math.pi / 3))
    ct = math.sqrt(1 - (math.sin(math.pi / 3) / 1.5) ** 2)
    f, _ = fresnel(m, (0, 0, 1), wi)
    assert abs(f - (r0 + (1 - r0) * (1 - ct ** 5))) < 1e-5
    # leaving the medium below the critical angle: cos(theta_i) term, ior 1/2.25
    wi = (math.sin(0.3), 0, -math.cos(0.3))
    f, ior = fresnel(m, (0, 0, 1), wi)
    assert abs(f - (r0 + (1 - r0) * (1 - math.cos(0.3) ** 5))) < 1e-5 and abs(ior - 1 / 2.25) < 1e-6
    # total internal reflection past asin(1/1.5)
    wi = (math.sin(0.8), 0, -math.cos(0.8))
    f, ior = fresnel(m, (0, 0, 1), wi)
    assert f == 1.0 and ior == -7          # factor left untouched


@pytest.mark.parametrize("theta,inside", [(0.2, False), (0.9, False), (0.3, True), (0.6, True)])
def test_refraction_obeys_snell(theta, inside):
    m, mp = mptr(mat(n=1.5))
    n = np.array([0, 0, 1.0])
    wi = np.array([math.sin(theta), 0, math.cos(theta)]) * (1, 1, -1 if inside else 1)
    out = np.zeros(4, np.float32)
    L.or_refract(mp, f4p(v4(*n)), f4p(v4(*wi)), f4p(out))
    t = out[:3].astype(np.float64)
    n1, n2 = (1.5, 1.0) if inside else (1.0, 1.5)
    assert abs(np.linalg.norm(t) - 1) < 1e-6
    assert np.sign(t[2]) == (1 if inside else -1)                      # crosses the surface
    assert abs(n1 * math.sin(theta) - n2 * math.hypot(t[0], t[1])) < 1e-5
    assert abs(t[1]) < 1e-7 and t[0] * wi[0] < 0                       # plane of incidence, other side


# ---------------------------------------------------------------- lights
def quad_light(pos, el, ew, normal, ke):
    return O.LightQuad.from_buffer_copy(bytes(Y.light_quad(pos, normal, ke, el, ew)))


def test_light_sampling_estimates_irradiance():
    """Mean of ke cos / pdf over light samples equals the irradiance
    integral ke * int cos cos_l / r^2 dA (scipy quadrature)."""
    Lq = quad_light((-0.25, 1, -0.4), (0.5, 0, 0), (0, 0, 0.6), (0, -1, 0), (10, 10, 10))
    p, n = np.array([0.1, 0, 0.05]), np.array([0.0, 1.0, 0.0])

    def integrand(z, x):
        r = np.array([x, 1, z]) - p
        d2 = r @ r
        return (r @ n) * r[1] / d2 ** 2        # cos * cos_l / r^2, with |r| folded in

    E = 10 * integrate.dblquad(integrand, -0.25, 0.25, -0.4, 0.2)[0]
    seed = C.c_uint32(4242)
    pdf = C.c_float()
    w = np.zeros(4, np.float32)
    acc, k = 0.0, 40000
    for _ in range(k):
        j = L.or_sample_lights(C.byref(Lq), f4p(v4(*p, 1)), f4p(v4(*n)), C.byref(seed), C.byref(pdf), f4p(w))
        assert j == 0
        q = p + w[:3]                           # the sampled point lies on the quad
        assert abs(q[1] - 1) < 1e-5 and -0.25 - 1e-5 <= q[0] <= 0.25 + 1e-5 and -0.4 - 1e-5 <= q[2] <= 0.2 + 1e-5
        acc += 10 * max(unit(w[:3]) @ n, 0) / pdf.value
    assert abs(acc / k - E) < 0.01 * E


def test_light_facing_away_gives_no_sample():
    Lq = quad_light((-0.25, 1, -0.25), (0.5, 0, 0), (0, 0, 0.5), (0, 1, 0), (10, 10, 10))
    seed = C.c_uint32(1)
    pdf = C.c_float()
    w = np.zeros(4, np.float32)
    assert L.or_sample_lights(C.byref(Lq), f4p(v4(0, 0, 0, 1)), f4p(v4(0, 1, 0)), C.byref(seed), C.byref(pdf),
                              f4p(w)) == -1


# ---------------------------------------------------------------- tonemap
def test_tonemap_is_gamma_with_unit_white():
    """With lum_white = 1 Reinhard's extended operator is the identity, so
    the output is the input to the power 1/2.2."""
    rng = np.random.default_rng(7)
    img = rng.uniform(0.01, 3, (100, 4)).astype(np.float32)
    out = O.tonemap(img)
    assert np.allclose(out, img.astype(np.float64) ** (1 / 2.2), rtol=3e-6)


def test_tonemap_hable_closed_form():
    """John Hable's curve (tonemap.cl:49-58), (x(Ax+CB)+DE)/(x(Ax+B)+DF) - E/F
    with A..F = .15 .5 .1 .2 .02 .3, then gamma: zero at zero, monotonic,
    tending to 1 - E/F = 14/15 for large x; values against float64."""
    A, B, Cc, D, E, F = 0.15, 0.50, 0.10, 0.20, 0.02, 0.30
    x = np.concatenate([[0.0, 0.5, 1.0, 11.2, 1e6], np.linspace(0.01, 50, 395)]).astype(np.float32)
    img = np.stack([x, x, x, x], -1)
    out = O.tonemap(img, 1)
    xd = x.astype(np.float64)
    curve = (xd * (A * xd + Cc * B) + D * E) / (xd * (A * xd + B) + D * F) - E / F
    exp = np.sign(curve) * np.abs(curve) ** (1 / 2.2)
    ok = curve > 1e-4
    assert np.allclose(out[ok, 0], exp[ok], rtol=2e-5)
    # at zero the float curve leaves one rounding residue, (DE)/(DF) - E/F =
    # 2^-27, whose 1/2.2 power is ~2.0e-4
    f = np.float32
    z = (f(D) * f(E)) / (f(D) * f(F)) - f(E) / f(F)
    assert z == f(2.0 ** -27) and abs(out[0, 0] - float(z) ** (1 / 2.2)) < 1e-8
    assert (np.diff(out[5:, 0]) > 0).all()
    assert abs(out[4, 0] - (14 / 15) ** (1 / 2.2)) < 1e-5
    # the four channels go through the same curve
    assert np.array_equal(out[:, 0].view(np.int32), out[:, 3].view(np.int32))
    # and differ from the shipped Reinhard operator (the identity before gamma)
    assert not np.allclose(out[1:4, 0], O.tonemap(img)[1:4, 0], rtol=1e-2)


# ---------------------------------------------------------------- whole-path checks
def box_scene(a, e, light_ke=(0, 0, 0)):
    """Closed cube [-1,1]^3 (12 triangles, normals inward), one grey
    material of albedo a and emission e; the light quad sits outside."""
    tris = []
    corners = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], float)
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    for f in faces:
        q = corners[list(f)]
        nrm = -unit(q.mean(0))
        for a_, b_, c_ in ((0, 1, 2), (0, 2, 3)):
            tris.append(tri(q[a_], q[b_], q[c_], nrm))
    tris = np.array(tris, O.TRI_DTYPE)
    mats = np.array([mat(kd=(a, a, a, 1), ke=(e, e, e, 1))], O.MAT_DTYPE)
    Lq = quad_light((5, 5, 5), (0.1, 0, 0), (0, 0, 0.1), (0, 1, 0), light_ke)
    return O.OScene(tris, mats, O.bvh_build(tris), Lq, heap_size=0)


def test_furnace_closed_form():
    """Every path in an emissive grey box returns the same value.  The
    reference adds a hit's emission twice per bounce (once when the bounce
    ray finds it, once inside evaluateDirectLighting), so B bounces give
    e * (1 + (1 + a) (1 - a^B) / (1 - a))."""
    a, e, B = 0.5, 0.2, 3
    s = box_scene(a, e)
    cam = O.camera((1, 0, 0), (0, 1, 0), (0, 0, -1), (0.1, 0.2, 0.3), 60.0)
    p = O.params(max_bounces=B, rr_threshold=100)
    img, _, _, _ = s.render(cam, 16, 16, p, O.mt19937(1, 2))
    exp = e * (1 + (1 + a) * (1 - a ** B) / (1 - a))
    assert np.allclose(img[..., :3], exp, rtol=2e-5)
    assert (img[..., 3] == 2).all()
    # gi off: emission only
    img, _, _, _ = s.render(cam, 8, 8, O.params(gi=0), O.mt19937(1, 1))
    assert np.allclose(img[..., :3], e, rtol=1e-6)


def test_russian_roulette_is_unbiased():
    """With roulette after bounce 2 and no bounce limit the mean is the
    infinite series e * (1 + (1 + a) / (1 - a))."""
    a, e = 0.5, 0.2
    s = box_scene(a, e)
    cam = O.camera((1, 0, 0), (0, 1, 0), (0, 0, -1), (0.1, 0.2, 0.3), 60.0)
    img, _, _, _ = s.render(cam, 64, 64, O.params(rr_threshold=2), O.mt19937(3, 4), threads=4)
    exp = e * (1 + (1 + a) / (1 - a))
    assert abs(img[..., 0].mean() - exp) < 0.01 * exp


def floor_scene():
    tris = np.array([tri((-5, 0, -5), (5, 0, 5), (5, 0, -5), (0, 1, 0)),
                     tri((-5, 0, -5), (-5, 0, 5), (5, 0, 5), (0, 1, 0))], O.TRI_DTYPE)
    mats = np.array([mat(kd=(0.5, 0.5, 0.5, 1))], O.MAT_DTYPE)
    Lq = quad_light((-0.75, 1, -0.75), (1.5, 0, 0), (0, 0, 1.5), (0, -1, 0), (10, 10, 10))
    return O.OScene(tris, mats, O.bvh_build(tris), Lq, heap_size=0)


def reference_mis_expectation(kd, ke, half):
    """Expected direct light at the origin (normal +y) under a square
    emitter of side 2*half at height 1, for the estimator of
    evaluateDirectLighting with MIS (udpt.cl:570-607): the light sample is
    weighted with the power heuristic of (pl, pb) at its own direction, but
    the BRDF sample's weight reuses pl of the *light sample* (an independent
    point on the emitter).  Midpoint quadrature; the weights of the two
    terms do not sum to one, so the result differs from kd/pi * E.

    E1 = f L  int_A cos cos_l / r^2 * pl^2 / (pl^2 + pb^2) dA
    E2 = f L  int_A int_A cos(x) cos_l(x) / r(x)^2 * pb(x)^2 / (pl(y)^2 + pb(x)^2) dA(x) dA(y) / A
    """
    n = 64
    g = (np.arange(n) + 0.5) / n * 2 * half - half
    x, z = np.meshgrid(g, g)
    r2 = x * x + 1 + z * z
    cos = 1 / np.sqrt(r2)                       # at the receiver and at the emitter alike
    area = (2 * half) ** 2
    dA = area / n / n
    pl = (r2 / cos / area).ravel()              # solid-angle pdf of light sampling
    pb = (cos / math.pi).ravel()                # cosine-hemisphere pdf
    geo = (cos * cos / r2).ravel()
    f = kd / math.pi
    e1 = f * ke * np.sum(geo * pl ** 2 / (pl ** 2 + pb ** 2)) * dA
    e2 = f * ke * np.sum(geo[:, None] * pb[:, None] ** 2 / (pl[None, :] ** 2 + pb[:, None] ** 2)) * dA / n / n
    return e1 + e2, f * ke * np.sum(geo) * dA


def test_mis_expectation_quadrature_is_consistent():
    """With the proper light pdf at the BRDF sample's own direction the two
    weights would sum to one; the quadrature reproduces that limit (so the
    shortfall it reports for the reference's weights is the weights', not
    the quadrature's)."""
    tot, plain = reference_mis_expectation(0.5, 10.0, 0.75)
    E = 10 * integrate.dblquad(lambda z, x: 1 / (x * x + 1 + z * z) ** 2, -0.75, 0.75, -0.75, 0.75)[0]
    assert abs(plain - 0.5 / math.pi * E) < 1e-3 * plain
    assert 0.9 < tot / plain < 0.99


@pytest.mark.parametrize("mis,fix", [(0, 0), (1, 0), (1, 1)])
def test_direct_light_converges_to_irradiance(mis, fix):
    """Direct light only (gi 0) at the floor point under the emitter, seen
    through a very narrow camera.  Without MIS the mean is kd/pi * E (E by
    quadrature); with MIS it is the reference estimator's own expectation
    (reference_mis_expectation).  The emitter is large enough that light
    and BRDF pdfs are of one order, so a wrong pdf or weight shows."""
    s = floor_scene()
    cam = O.camera((1, 0, 0), unit((0, 1, -1)), unit((0, -1, -1)), (0, 0.5, 0.5), 0.01)
    img, _, _, _ = s.render(cam, 64, 64, O.params(gi=0, mis=mis, fix_mis=fix), O.mt19937(9, 16), threads=8)
    tot, plain = reference_mis_expectation(0.5, 10.0, 0.75)
    exp = tot if mis else plain
    assert abs(img[..., 0].mean() - exp) < 0.006 * exp


def test_trace_matches_brute_force():
    """Closest hit through the oracle's BVH (no heap cap) equals the
    minimum over every triangle; the 1500-entry cap only changes rays it
    flags as truncated."""
    sc = Y.gen_icosphere_room(11, 6, 3)
    nodes = O.bvh_build(sc.tris)
    far_light = quad_light((50, 50, 50), (0.1, 0, 0), (0, 0, 0.1), (0, 1, 0), (1, 1, 1))
    exact = O.OScene(sc.tris, sc.mats, nodes, far_light, heap_size=0)
    capped = O.OScene(sc.tris, sc.mats, nodes, far_light, heap_size=1500)
    brute = O.OScene(sc.tris, sc.mats, None, far_light, heap_size=0)
    rng = np.random.default_rng(8)
    n = 300
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = rng.uniform(-0.9, 0.9, (n, 3)) * (sc.tris["v1"][:, :3].max(0) - sc.tris["v1"][:, :3].min(0)) / 2 \
        + (sc.tris["v1"][:, :3].max(0) + sc.tris["v1"][:, :3].min(0)) / 2
    d = rng.normal(size=(n, 3))
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 3] = np.inf
    t_e, _, tt_e, _ = exact.trace(rays)
    t_b, _, tt_b, _ = brute.trace(rays)
    assert np.array_equal(t_e, t_b) and np.array_equal(tt_e, tt_b)
    for k in range(0, 60):                          # spot check against or_ray_tri one by one
        best, bt = -1, np.inf
        for j in range(len(sc.tris)):
            g, tl = ray_tri(sc.tris[j], rays[k, :3], rays[k, 4:7], bt)
            if g:
                best, bt = j, tl
        assert best == t_e[k] and (best < 0 or bt == tt_e[k])
    t_c, _, tt_c, fl = capped.trace(rays)
    ok = (fl & 1) == 0
    assert np.array_equal(t_c[ok], t_e[ok])


def test_bdpt_furnace_closed_form():
    """bdpt.cl shading in an emissive grey box.  The light quad emits but
    sits outside the box facing away, so no light subpath vertex reaches it
    (lp_len = 1: no connections) and NEE from inside returns the emission
    only.  Every eye vertex i adds T_i (ke + evaluateDirectLighting) w_i
    with T_i = a^i, w_i = 0.1^i (ks floor 0.1, bdpt.cl:592), so B eye
    vertices give 2e (1 - (0.1 a)^B) / (1 - 0.1 a).  Light vertices and
    connections are pinned in tests/test_oracle_bdpt.py."""
    a, e, B = 0.5, 0.2, 5
    s = box_scene(a, e, light_ke=(5, 5, 5))
    cam = O.camera((1, 0, 0), (0, 1, 0), (0, 0, -1), (0.1, 0.2, 0.3), 60.0)
    p = O.params(1, bdpt_bounces=B, rr_threshold=100)
    img, _, _, _ = s.render(cam, 16, 16, p, O.mt19937(2, 2))
    exp = 2 * e * (1 - (0.1 * a) ** B) / (1 - 0.1 * a)
    assert np.allclose(img[..., :3], exp, rtol=2e-5)


def test_heap_size_truncation_is_counted():
    """The reference's breadth-first list stops at HEAP_SIZE entries
    (udpt.cl:8, 288-324); the oracle counts such queries.  A list of 3
    entries truncates most queries of the Cornell box; 1500 none."""
    s = Y.gen_cornell()
    cam = O.Cam.from_buffer_copy(bytes(s.cam))
    counts = {}
    for heap in (3, 1500):
        osc = O.OScene(s.tris, s.mats, s.bvh(), s.light, heap_size=heap)
        _, _, rays, _ = osc.render(cam, 16, 16, O.params(0, mis=1, max_bounces=8), O.mt19937(1, 1))
        counts[heap] = (osc.last_counts["truncated"], rays)
    assert counts[3][0] > 0.5 * counts[3][1] and counts[1500][0] == 0, counts
