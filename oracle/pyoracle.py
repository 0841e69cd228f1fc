"""ctypes wrapper of the CPU oracle (oracle.c) and of the reference's own
BVH builder compiled into oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke()
and by bench.py's CPU-baseline leg, always as the checker (or the timed CPU
baseline), never as a product path.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "liboracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libref_bvh.so")

# layouts identical to include/yune_types.h
TRI_DTYPE = np.dtype(
    [("v1", "<f4", 4), ("v2", "<f4", 4), ("v3", "<f4", 4), ("vn1", "<f4", 4), ("vn2", "<f4", 4),
     ("vn3", "<f4", 4), ("matID", "<i4"), ("pad", "<f4", 3)], align=True)
MAT_DTYPE = np.dtype(
    [("ke", "<f4", 4), ("kd", "<f4", 4), ("ks", "<f4", 4), ("n", "<f4"), ("k", "<f4"), ("px", "<f4"),
     ("py", "<f4"), ("alpha_x", "<f4"), ("alpha_y", "<f4"), ("is_specular", "<i4"),
     ("is_transmissive", "<i4")], align=True)
NODE_DTYPE = np.dtype(
    [("p_min", "<f4", 4), ("p_max", "<f4", 4), ("vert_list", "<i4", 10), ("child_idx", "<i4"),
     ("vert_len", "<i4")], align=True)

F4 = C.c_float * 4


class Cam(C.Structure):
    _fields_ = [("r1", F4), ("r2", F4), ("r3", F4), ("r4", F4), ("view_plane_dist", C.c_float),
                ("pad", C.c_float * 3)]


class LightQuad(C.Structure):
    _fields_ = [("pos", F4), ("normal", F4), ("ke", F4), ("edge_l", F4), ("edge_w", F4)]


class RenderParams(C.Structure):
    _fields_ = [("mode", C.c_int32), ("gi", C.c_int32), ("mis", C.c_int32), ("max_bounces", C.c_int32),
                ("rr_threshold", C.c_int32), ("bdpt_bounces", C.c_int32), ("oren_nayar", C.c_int32),
                ("fix_mis", C.c_int32), ("background", C.c_float * 4)]


class Scene(C.Structure):
    # yn_light_quad is 16-byte aligned in C: explicit pad before it
    _fields_ = [("tris", C.c_void_p), ("n_tri", C.c_int), ("mats", C.c_void_p), ("n_mat", C.c_int),
                ("nodes", C.c_void_p), ("n_node", C.c_int), ("_pad0", C.c_int), ("light", LightQuad),
                ("heap_size", C.c_int), ("_pad1", C.c_int * 3)]


assert Scene.light.offset == 48 and Scene.heap_size.offset == 128 and C.sizeof(Scene) == 144


# or_pvertex (oracle.h): a BDPT subpath vertex
PVERTEX_DTYPE = np.dtype([("p", "<f4", 4), ("n", "<f4", 4), ("dir", "<f4", 4), ("contrib", "<f4", 4),
                          ("tri", "<i4"), ("light", "<i4"), ("pad", "<i4", 2)])
assert PVERTEX_DTYPE.itemsize == 80


class Capture(C.Structure):
    _fields_ = [("prim_tri", C.c_void_p), ("prim_light", C.c_void_p), ("prim_t", C.c_void_p),
                ("shadow_ray", C.c_void_p), ("shadow_occ", C.c_void_p), ("trunc", C.c_void_p)]


_o = None
_r = None


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def load(path: str) -> C.CDLL:
    """The oracle library at `path` with its signatures declared."""
    L = C.CDLL(path)
    L.or_wang_hash.restype = C.c_uint32
    L.or_wang_hash.argtypes = [C.c_uint32]
    L.or_xor_shift.restype = C.c_uint32
    L.or_xor_shift.argtypes = [C.c_uint32]
    L.or_u2f.restype = C.c_float
    L.or_u2f.argtypes = [C.c_uint32]
    L.or_pixel_seed.restype = C.c_uint32
    L.or_pixel_seed.argtypes = [C.c_uint32, C.c_int, C.c_int, C.c_int]
    L.or_mt19937.argtypes = [C.c_uint32, C.c_int, C.c_void_p]
    L.or_camera.argtypes = [C.c_void_p] * 4 + [C.c_float, C.POINTER(Cam)]
    L.or_bvh_build.restype = C.c_int
    L.or_bvh_build.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int]
    L.or_trace.argtypes = [C.POINTER(Scene), C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                           C.c_void_p]
    L.or_render.restype = C.c_int
    L.or_render.argtypes = [C.POINTER(Scene), C.POINTER(Cam), C.c_int, C.c_int, C.POINTER(RenderParams),
                            C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                            C.c_void_p, C.POINTER(Capture), C.c_void_p]
    L.or_sample.argtypes = [C.POINTER(Scene), C.POINTER(Cam), C.c_int, C.c_int, C.POINTER(RenderParams),
                            C.c_uint32, C.c_int, C.c_int, C.c_void_p]
    L.or_tonemap.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    L.or_tonemap_hable.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    L.or_cosine_sample.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32), C.c_void_p,
                                   C.POINTER(C.c_float)]
    L.or_phong_sample.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32),
                                  C.c_void_p, C.POINTER(C.c_float)]
    L.or_eval_brdf.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_float, C.c_int,
                               C.c_void_p]
    L.or_fresnel.restype = C.c_float
    L.or_fresnel.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]
    L.or_refract.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.or_sample_lights.restype = C.c_int
    L.or_sample_lights.argtypes = [C.POINTER(LightQuad), C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_float), C.c_void_p]
    L.or_ray_box.restype = C.c_int
    L.or_ray_box.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.or_primary_ray.argtypes = [C.POINTER(Cam), C.c_int, C.c_int, C.c_uint32, C.c_int, C.c_int, C.c_void_p,
                                 C.c_void_p]
    L.or_phong_pdf.restype = C.c_float
    L.or_phong_pdf.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    for fn in ("or_fm_sinf", "or_fm_cosf"):
        getattr(L, fn).restype = C.c_float
        getattr(L, fn).argtypes = [C.c_float]
    L.or_fm_powf.restype = C.c_float
    L.or_fm_powf.argtypes = [C.c_float, C.c_float]
    L.or_ray_tri.restype = C.c_int
    L.or_ray_tri.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]
    L.or_bdpt_light_path.restype = C.c_int
    L.or_bdpt_light_path.argtypes = [C.POINTER(Scene), C.POINTER(RenderParams), C.POINTER(C.c_uint32),
                                     C.c_void_p]
    L.or_bdpt_eye_path.restype = C.c_int
    L.or_bdpt_eye_path.argtypes = [C.POINTER(Scene), C.POINTER(RenderParams), C.c_void_p, C.c_void_p,
                                   C.POINTER(C.c_uint32), C.c_void_p]
    L.or_bdpt_combine.argtypes = [C.POINTER(Scene), C.POINTER(RenderParams), C.c_void_p, C.c_int, C.c_void_p,
                                  C.c_int, C.POINTER(C.c_uint32), C.c_void_p, C.c_void_p]
    return L


def lib():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            raise ImportError(f"{ORACLE_SO} not built (make -C oracle)")
        _o = load(ORACLE_SO)
    return _o


@contextlib.contextmanager
def variant(path: str):
    """Within the block, every oracle call goes to the library at `path`
    (a build with a planted mistake, for tests that show the pins catch
    it)."""
    global _o
    saved = lib()
    _o = load(path)
    try:
        yield _o
    finally:
        _o = saved


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_bvh_build(tris: np.ndarray, bins: int = 20) -> np.ndarray:
    """Nodes from the reference's src/BVH.cpp itself (oracle/_ref)."""
    global _r
    if _r is None:
        _r = C.CDLL(REF_SO)
        _r.ref_bvh_build.restype = C.c_int
        _r.ref_bvh_build.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int]
    tris = np.ascontiguousarray(tris, dtype=TRI_DTYPE)
    out = np.zeros(4 * len(tris) + 64, dtype=NODE_DTYPE)
    n = _r.ref_bvh_build(_p(tris), len(tris), bins, _p(out), len(out))
    if n > len(out):
        out = np.zeros(n, dtype=NODE_DTYPE)
        n = _r.ref_bvh_build(_p(tris), len(tris), bins, _p(out), len(out))
    return out[:n]


def bvh_build(tris: np.ndarray, bins: int = 20) -> np.ndarray:
    tris = np.ascontiguousarray(tris, dtype=TRI_DTYPE)
    out = np.zeros(4 * len(tris) + 64, dtype=NODE_DTYPE)
    n = lib().or_bvh_build(_p(tris), len(tris), bins, _p(out), len(out))
    if n < 0:
        raise ValueError("oracle BVH build failed")
    if n > len(out):
        out = np.zeros(n, dtype=NODE_DTYPE)
        lib().or_bvh_build(_p(tris), len(tris), bins, _p(out), len(out))
    return out[:n]


def mt19937(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.uint32)
    lib().or_mt19937(seed, n, _p(out))
    return out


def camera(side, up, look_at, eye, yfov=60.0) -> Cam:
    arrs = [np.ascontiguousarray(v, np.float32) for v in (side, up, look_at, eye)]
    c = Cam()
    lib().or_camera(*[_p(a) for a in arrs], float(yfov), C.byref(c))
    return c


class OScene:
    """Keeps the arrays alive behind an or_scene."""

    def __init__(self, tris, mats, nodes, light, heap_size=1500):
        self.tris = np.ascontiguousarray(tris, dtype=TRI_DTYPE)
        self.mats = np.ascontiguousarray(mats, dtype=MAT_DTYPE)
        self.nodes = np.ascontiguousarray(nodes if nodes is not None else np.zeros(0, NODE_DTYPE),
                                          dtype=NODE_DTYPE)
        lq = LightQuad.from_buffer_copy(bytes(light))
        self.s = Scene(tris=_p(self.tris).value, n_tri=len(self.tris), mats=_p(self.mats).value,
                       n_mat=len(self.mats), nodes=_p(self.nodes).value if len(self.nodes) else None,
                       n_node=len(self.nodes), light=lq, heap_size=heap_size)

    def trace(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = len(rays)
        tri = np.empty(n, np.int32)
        light = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        flags = np.empty(n, np.int32)
        lib().or_trace(C.byref(self.s), _p(rays), n, _p(tri), _p(light), _p(t), _p(flags))
        return tri, light, t, flags

    def render(self, cam, W, H, params, seeds, accum=None, reset_first=True, rect=None, threads=1,
               capture=False):
        seeds = np.ascontiguousarray(seeds, np.uint32)
        if accum is None:
            accum = np.zeros((H, W, 4), np.float32)
        x0, y0, x1, y1 = rect if rect is not None else (0, 0, W, H)
        cap = None
        capst = None
        if capture:
            n = W * H
            cap = {"prim_tri": np.full(n, -9, np.int32), "prim_light": np.full(n, -9, np.int32),
                   "prim_t": np.zeros(n, np.float32), "shadow_ray": np.zeros((n, 8), np.float32),
                   "shadow_occ": np.full(n, -1, np.int32), "trunc": np.zeros(n, np.int32)}
            capst = Capture(*[_p(cap[k]).value for k in ("prim_tri", "prim_light", "prim_t", "shadow_ray",
                                                         "shadow_occ", "trunc")])
        rays = np.zeros(8, np.uint64)
        pc = C.byref(params) if not isinstance(params, C._Pointer) else params
        r = lib().or_render(C.byref(self.s), C.byref(cam), W, H, pc, _p(seeds), len(seeds), int(reset_first),
                            x0, y0, x1, y1, threads, _p(accum), C.byref(capst) if capst else None, _p(rays))
        if r != 0:
            raise ValueError("or_render failed")
        self.last_counts = {"truncated": int(rays[2]), "conn_traced": int(rays[3]), "conn_occluded": int(rays[4])}
        return accum, cap, int(rays[0]), int(rays[1])

    def bdpt_light_path(self, params, seed: int):
        """createLightPath from `seed`: (vertices, seed after)."""
        out = np.zeros(64, PVERTEX_DTYPE)
        sd = C.c_uint32(seed)
        n = lib().or_bdpt_light_path(C.byref(self.s), C.byref(params), C.byref(sd), _p(out))
        if n < 0:
            raise ValueError("bad BDPT parameters")
        return out[:n].copy(), sd.value

    def bdpt_eye_path(self, params, o, d, seed: int):
        """createEyePath from the camera ray (o, d): (vertices, seed after)."""
        out = np.zeros(64, PVERTEX_DTYPE)
        sd = C.c_uint32(seed)
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        n = lib().or_bdpt_eye_path(C.byref(self.s), C.byref(params), _p(o), _p(d), C.byref(sd), _p(out))
        if n < 0:
            raise ValueError("bad BDPT parameters")
        return out[:n].copy(), sd.value

    def bdpt_combine(self, params, lp, ep, seed: int):
        """The eye-vertex loop of bdpt.cl shading over given subpaths:
        (colour, seed after, connection rays traced, occluded)."""
        lp = np.ascontiguousarray(lp, PVERTEX_DTYPE)
        ep = np.ascontiguousarray(ep, PVERTEX_DTYPE)
        out = np.zeros(4, np.float32)
        cnt = np.zeros(2, np.uint64)
        sd = C.c_uint32(seed)
        lib().or_bdpt_combine(C.byref(self.s), C.byref(params), _p(lp), len(lp), _p(ep), len(ep), C.byref(sd),
                              _p(out), _p(cnt))
        return out, sd.value, int(cnt[0]), int(cnt[1])


def params(mode=0, **over) -> RenderParams:
    """udpt.cl (mode 0) / bdpt.cl (mode 1) defaults with overrides."""
    p = RenderParams(mode, 1, 0, 100000, 4 if mode == 1 else 6, 20, 0, 0)
    for i, v in enumerate((0.4, 0.4, 0.4, 1.0)):
        p.background[i] = v
    for k, v in over.items():
        if k == "background":
            for i, x in enumerate(v):
                p.background[i] = x
        else:
            setattr(p, k, v)
    return p


def tonemap(img: np.ndarray, op: int = 0) -> np.ndarray:
    """tonemap.cl: op 0 Reinhard (as shipped), op 1 John Hable's curve."""
    img = np.ascontiguousarray(img, np.float32)
    out = np.empty_like(img)
    (lib().or_tonemap_hable if op == 1 else lib().or_tonemap)(_p(img), img.size // 4, _p(out))
    return out
