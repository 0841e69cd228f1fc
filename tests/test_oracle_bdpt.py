"""Pins the oracle's bidirectional pieces against mathematics.

kernels/legacy/bdpt.cl builds a light subpath (createLightPath :432-511), an
eye subpath (createEyePath :513-560) and sums, per eye vertex i,
T_i (ke + direct light) w_i plus the connections to light vertices j >= 1,
G f_eye f_light contrib_j T_i, weighted w_i (1 - ks_i) (shading :562-640).
These tests fix each piece by a closed form or a quadrature, independent of
the oracle's own code:

* light vertex 0 is the quad point pos + edge_l r2 + edge_w r1 with
  contrib ke * A, r1 and r2 the next two xorshift draws (r1 first, as an FP64
  division); vertex 1, reached by a cosine-weighted ray, carries
  cos / pdf * contrib_0 = pi * ke * A; a diffuse vertex multiplies by kd;
* hand-placed subpaths: the connection sum is G = cos_e cos_l / d^2 times
  kd/pi at both ends, contrib_j and T_i, weights w_0 = 1, w_{i+1} = w_i ks_i
  with ks floored at 0.1, unoccluded connections only, emission counted
  twice (once directly, once inside evaluateDirectLighting);
* a two-plane scene whose light emits: the expected pixel value of the
  full estimator against a quadrature of the one-bounce integral.

tests/test_oracle_planted.py builds the oracle with planted mistakes and
shows these checks fail on each.
"""
import numpy as np
import pytest

from oracle import pyoracle as O

M32 = 0xFFFFFFFF


def xorshift(x):
    """xor_shift (udpt.cl:1141-1147), the 13/17/5 xorshift32."""
    x ^= (x << 13) & M32
    x ^= x >> 17
    x ^= (x << 5) & M32
    return x


def tri(a, b, c, n):
    t = np.zeros((), O.TRI_DTYPE)
    t["v1"], t["v2"], t["v3"] = (*a, 1), (*b, 1), (*c, 1)
    t["vn1"] = t["vn2"] = t["vn3"] = (*n, 0)
    return t


def mat(kd, ke=(0, 0, 0), ks=(0, 0, 0)):
    m = np.zeros((), O.MAT_DTYPE)
    m["ke"], m["kd"], m["ks"] = (*ke, 1), (*kd, 1), (*ks, 1)
    m["n"] = m["k"] = m["px"] = m["py"] = 1
    m["alpha_x"] = m["alpha_y"] = 100
    return m


def light(pos, el, ew, normal, ke):
    q = O.LightQuad()
    for name, v, w in (("pos", pos, 1.0), ("normal", normal, 0.0), ("ke", ke, 0.0), ("edge_l", el, 0.0),
                       ("edge_w", ew, 0.0)):
        arr = getattr(q, name)
        for i in range(3):
            arr[i] = float(v[i])
        arr[3] = w
    return q


# Two planes: a floor (y = 0, x in [-10, 10], z in [-1, 8]) and a wall band
# (z = -1, y in [2.5, 6]) above the light's plane, so the light (facing
# down at y = 2) reaches the floor only; seen from the band, the light quad
# hides no floor point.
KD_FLOOR = np.array([0.8, 0.7, 0.6])
KD_WALL = np.array([0.75, 0.8, 0.85])
KE = np.array([20.0, 16.0, 12.0])
LPOS, LEL, LEW = np.array([-0.25, 2.0, 5.0]), np.array([0.5, 0.0, 0.0]), np.array([0.0, 0.0, 0.5])
FLOOR_X, FLOOR_Z, BAND_Y = (-10.0, 10.0), (-1.0, 8.0), (2.5, 6.0)


def two_planes(light_ke=KE, wall_ke=(0, 0, 0), light_pos=LPOS):
    x0, x1 = FLOOR_X
    z0, z1 = FLOOR_Z
    y0, y1 = BAND_Y
    tris = np.array([tri((x0, 0, z0), (x1, 0, z1), (x1, 0, z0), (0, 1, 0)),
                     tri((x0, 0, z0), (x0, 0, z1), (x1, 0, z1), (0, 1, 0)),
                     tri((x0, y0, -1), (x1, y0, -1), (x1, y1, -1), (0, 0, 1)),
                     tri((x0, y0, -1), (x1, y1, -1), (x0, y1, -1), (0, 0, 1))], O.TRI_DTYPE)
    tris["matID"] = [0, 0, 1, 1]
    mats = np.array([mat(KD_FLOOR), mat(KD_WALL, ke=wall_ke)], O.MAT_DTYPE)
    Lq = light(light_pos, LEL, LEW, (0, -1, 0), light_ke)
    return O.OScene(tris, mats, O.bvh_build(tris), Lq, heap_size=0)


def bdpt(**kw):
    kw.setdefault("rr_threshold", 100)
    kw.setdefault("mis", 0)
    return O.params(1, **kw)


# ---------------------------------------------------------------- light subpath
def check_light_subpath(n_seeds=4000):
    """Vertex 0 on the quad with contrib ke A; vertex 1 (cosine emission)
    pi ke A on the floor along its ray; vertex 2 (diffuse floor) kd times
    vertex 1's contrib."""
    s = two_planes()
    p = bdpt(bdpt_bounces=3)
    A = np.linalg.norm(LEL) * np.linalg.norm(LEW)
    rng = np.random.default_rng(5)
    n1 = n2 = 0
    for seed in rng.integers(1, 1 << 32, n_seeds, dtype=np.uint64):
        seed = int(seed)
        lp, _ = s.bdpt_light_path(p, seed)
        s1 = xorshift(seed)
        s2 = xorshift(s1)
        r1, r2 = s1 / 4294967295.0, s2 / 4294967296.0
        v0 = lp[0]
        assert np.allclose(v0["p"][:3], LPOS + LEL * r2 + LEW * r1, atol=2e-6), "light vertex 0 position"
        assert np.allclose(v0["contrib"][:3], KE * A, rtol=1e-6), "light vertex 0 contrib"
        assert np.array_equal(v0["n"][:3], [0, -1, 0]) and v0["light"] == 0
        if len(lp) >= 2:
            v1 = lp[1]
            n1 += 1
            assert abs(v1["p"][1]) < 1e-5 and v1["tri"] in (0, 1), "vertex 1 lies on the floor"
            d = v1["dir"][:3].astype(np.float64)
            assert abs(np.linalg.norm(d) - 1) < 1e-6 and d[1] < 0
            # the vertex lies on the ray from vertex 0
            off = v1["p"][:3] - v0["p"][:3]
            assert np.linalg.norm(np.cross(off, d)) < 1e-4 * max(1.0, np.linalg.norm(off))
            assert np.allclose(v1["contrib"][:3], np.pi * KE * A, rtol=3e-5), "light vertex 1 contrib"
        if len(lp) >= 3:
            n2 += 1
            v2 = lp[2]
            assert v2["tri"] in (2, 3), "vertex 2 lies on the wall band"
            assert np.allclose(v2["contrib"][:3], lp[1]["contrib"][:3] * KD_FLOOR, rtol=3e-5), "vertex 2 contrib"
    assert n1 > 0.5 * n_seeds and n2 > 20, (n1, n2)


def test_light_subpath_vertices():
    check_light_subpath()


def check_eye_subpath(n_seeds=2000):
    """Camera vertex contrib 1; a diffuse bounce multiplies by kd."""
    s = two_planes()
    p = bdpt(bdpt_bounces=3)
    o = np.array([0.3, 3.4, 12.0, 1.0], np.float32)
    rng = np.random.default_rng(6)
    n1 = 0
    for seed in rng.integers(1, 1 << 32, n_seeds, dtype=np.uint64):
        d = np.array([0.01, 0.02, -1.0, 0.0])
        d[:3] /= np.linalg.norm(d[:3])
        ep, _ = s.bdpt_eye_path(p, o, d.astype(np.float32), int(seed))
        assert len(ep) >= 1 and ep[0]["tri"] in (2, 3)
        assert np.array_equal(ep[0]["contrib"], [1, 1, 1, 1])
        if len(ep) >= 2:
            n1 += 1
            assert ep[1]["tri"] in (0, 1)
            assert np.allclose(ep[1]["contrib"][:3], KD_WALL, rtol=3e-5), "eye vertex 1 contrib"
        if len(ep) >= 3:
            assert np.allclose(ep[2]["contrib"][:3], KD_WALL * KD_FLOOR, rtol=5e-5), "eye vertex 2 contrib"
    assert n1 > 0.1 * n_seeds


def test_eye_subpath_vertices():
    check_eye_subpath()


# ---------------------------------------------------------------- connections
def vertex(p, n, d, contrib, tri_id):
    v = np.zeros((), O.PVERTEX_DTYPE)
    v["p"], v["n"] = (*p, 1), (*n, 0)
    dd = np.asarray(d, np.float64)
    v["dir"] = (*(dd / np.linalg.norm(dd)), 0)
    v["contrib"] = (*contrib, 1)
    v["tri"], v["light"] = tri_id, -1
    return v


def geometry(e, le):
    """G of bdpt.cl:613 between vertices e and l: cos at both ends / d^2."""
    dv = le["p"][:3].astype(np.float64) - e["p"][:3].astype(np.float64)
    d2 = dv @ dv
    w = dv / np.sqrt(d2)
    return max(w @ e["n"][:3], 0.0) * max(-w @ le["n"][:3], 0.0) / d2


def check_connection_sum():
    """Two eye vertices and three light vertices placed by hand.  The
    scene's light emits nothing and lies far away, so evaluateDirectLighting
    returns the vertex's emission only; the sum is then

      T_0 2 ke_wall w_0 + w_0 (1 - ks_0) C_1 G(e0, l1) kd_w/pi kd_f/pi T_0
      + T_1 2 ke_floor w_1 + w_1 (1 - ks_1) C_2 G(e1, l2) kd_f/pi kd_w/pi T_1

    with w_0 = 1, ks = 0.1 (diffuse, floored), w_1 = 0.1.  e0-l2 and e0-l3
    face away from each other and e1-l1 lie in one plane (no ray); e1-l3 is
    a ray the wall band blocks."""
    ke_wall = np.array([0.05, 0.04, 0.03])
    s = two_planes(light_ke=(0, 0, 0), wall_ke=ke_wall, light_pos=(100, 100, 100))
    p = bdpt(bdpt_bounces=4)
    cam = np.array([0.2, 3.3, 12.0])
    e0 = vertex((0.3, 3.2, -1.0), (0, 0, 1), (0.3, 3.2, -1.0) - cam, (1, 1, 1), 2)
    T1 = np.array([0.5, 0.6, 0.7])
    e1 = vertex((-0.5, 0.0, 1.5), (0, 1, 0), (-0.8, -3.2, 2.5), T1, 0)
    C1, C2, C3 = np.array([3.0, 2.0, 1.0]), np.array([1.5, 2.5, 0.5]), np.array([9.0, 9.0, 9.0])
    l0 = vertex((0, 2, 5), (0, -1, 0), (0, -1, 0), (5, 5, 5), -1)
    l0["light"] = 0
    l1 = vertex((1.0, 0.0005, 0.4), (0, 1, 0), (0.2, -1, -0.3), C1, 0)
    l2 = vertex((-0.7, 4.0, -0.9995), (0, 0, 1), (0.1, 0.3, -1), C2, 2)
    l3 = vertex((0.4, 4.0, -1.5), (0, 0, 1), (0, 0, -1), C3, 2)
    lp = np.array([l0, l1, l2, l3])
    ep = np.array([e0, e1])
    got, _, traced, occluded = s.bdpt_combine(p, lp, ep, 12345)
    fw, ff = KD_WALL / np.pi, KD_FLOOR / np.pi
    w0, w1 = 1.0, 0.1
    # the floor (e1's material) emits nothing
    exp = (2 * ke_wall * w0 + w0 * 0.9 * C1 * geometry(e0, l1) * fw * ff
           + w1 * 0.9 * C2 * geometry(e1, l2) * ff * fw * T1)
    assert (traced, occluded) == (3, 1), (traced, occluded)
    assert np.allclose(got[:3], exp, rtol=2e-5), (got, exp)


def test_connection_sum_at_hand_placed_vertices():
    check_connection_sum()


# ---------------------------------------------------------------- quadrature
CAM_EYE = np.array([0.0, 3.5, 12.0])
FOV = 2.0


def gauss(a, b, n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (b - a) * x + 0.5 * (b + a), 0.5 * (b - a) * w


def one_bounce_integral(xs, refine=1):
    """I(x) = kd_w/pi * int_floor L1(y) cos_x cos_y / r^2 dA_y, with L1 the
    floor's radiance after one reflection of the quad's emission:
    L1 = kd_f/pi * ke * int_quad cos_0 cos_1 / r^2 dA_0 (both planes
    horizontal: cos_0 = cos_1 = h / r)."""
    # floor grid: composite Gauss-Legendre, finer around the light's footprint
    def axis(lo, hi, c, n_in, n_out):
        edges = np.unique(np.concatenate([np.linspace(lo, hi, n_out), np.linspace(c - 3, c + 3, n_in)]))
        edges = edges[(edges >= lo) & (edges <= hi)]
        pts, wts = [], []
        for a, b in zip(edges[:-1], edges[1:]):
            x, w = gauss(a, b, 4)
            pts.append(x)
            wts.append(w)
        return np.concatenate(pts), np.concatenate(wts)
    cx, cz = (LPOS + 0.5 * (LEL + LEW))[[0, 2]]
    fx, wx = axis(*FLOOR_X, cx, 60 * refine + 1, 40 * refine + 1)
    fz, wz = axis(*FLOOR_Z, cz, 60 * refine + 1, 18 * refine + 1)
    FX, FZ = np.meshgrid(fx, fz, indexing="ij")
    WA = np.outer(wx, wz)
    qx, qwx = gauss(LPOS[0], LPOS[0] + LEL[0], 12 * refine)
    qz, qwz = gauss(LPOS[2], LPOS[2] + LEW[2], 12 * refine)
    h = LPOS[1]
    E = np.zeros_like(FX)
    for x0, w0 in zip(qx, qwx):
        for z0, w1 in zip(qz, qwz):
            r2 = (FX - x0) ** 2 + (FZ - z0) ** 2 + h * h
            E += w0 * w1 * h * h / (r2 * r2)
    out = []
    for x in xs:
        dx, dy, dz = FX - x[0], -x[1], FZ - x[2]
        r2 = dx * dx + dy * dy + dz * dz
        cos_x = dz / np.sqrt(r2)                # wall normal +z
        cos_y = x[1] / np.sqrt(r2)              # floor normal +y
        J = (E * np.maximum(cos_x, 0) * cos_y / r2 * WA).sum()
        out.append(KD_WALL / np.pi * KD_FLOOR / np.pi * KE * J)
    return np.mean(out, 0)


def view_points(n=5):
    """Wall points across the camera's view (pinhole at CAM_EYE looking
    down -z, vertical FOV 2 degrees read with pi = 3.14 as Camera.cpp:49)."""
    half = (12.0 + 1.0) * np.tan(FOV * 3.14 / 360.0)
    g = np.linspace(-half, half, n)
    return [np.array([CAM_EYE[0] + a, CAM_EYE[1] + b, -1.0]) for a in g for b in g]


def check_two_plane_quadrature(bounces, W=64, H=64, F=32):
    """The full estimator at wall pixels, against the integral.  The light
    cannot reach the band (NEE there is zero), so with 1 eye vertex the
    pixel is the connections' share, w_0 (1 - ks_0) = 0.9 of I, and with 2
    the eye bounce's NEE on the floor adds w_1 = 0.1 of I.  A connection ray
    runs to the light vertex's exact distance on the floor, so the floor can
    occlude it by rounding: the connections' share is bounded by 0 (every
    ray occluded) and 0.9 I (none), and expected at 0.9 (1 - r) I for the
    oracle's measured occluded fraction r."""
    s = two_planes()
    cam = O.camera((1, 0, 0), (0, 1, 0), (0, 0, -1), CAM_EYE, FOV)
    p = bdpt(bdpt_bounces=bounces)
    img, _, _, _ = s.render(cam, W, H, p, O.mt19937(11, F), threads=8)
    r = s.last_counts["conn_occluded"] / max(s.last_counts["conn_traced"], 1)
    I = one_bounce_integral(view_points())
    pix = img[..., :3].reshape(-1, 3).astype(np.float64)
    mean = pix.mean(0)
    sem = pix.std(0) / np.sqrt(len(pix))
    eye = 0.1 if bounces >= 2 else 0.0
    lo, hi = eye * I, (eye + 0.9) * I
    exp = (eye + 0.9 * (1 - r)) * I
    assert (mean > lo - 4 * sem).all() and (mean < hi + 4 * sem).all(), (mean, lo, hi)
    tol = 4 * sem + 0.01 * exp
    assert (np.abs(mean - exp) < tol).all(), (mean, exp, sem, r)
    return r, mean, exp


@pytest.mark.parametrize("bounces", [1, 2])
def test_two_plane_connections_against_quadrature(bounces):
    r, mean, exp = check_two_plane_quadrature(bounces)
    print(f"bdpt_bounces {bounces}: connection rays occluded by rounding {r:.3f}, pixel mean {mean}, expected {exp}")
    assert 0.0 < r < 1.0


def test_quadrature_converges():
    """The integral the checks above use does not move when its cells and
    the quad's Gauss points double, nor with 7 x 7 view points instead of
    5 x 5."""
    a = one_bounce_integral(view_points())
    b = one_bounce_integral(view_points(), refine=2)
    c = one_bounce_integral(view_points(7))
    assert np.allclose(a, b, rtol=1e-4), (a, b)
    assert np.allclose(a, c, rtol=1e-3), (a, c)
