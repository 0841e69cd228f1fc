"""yn_trace_noalloc: the allocation-free variant of yn_trace (no counterpart
in the reference; for callers that cannot allocate device memory).

Rays, triangle ids, light ids and lengths live in device buffers the test
allocates itself through the HIP runtime (ctypes).  Two calls in a row (the
second with fewer rays, so the static scratch array still holds the first
call's contents) must equal yn_trace on the same rays bit for bit:
triangle id, light id and the bits of t, for closest-hit and any-hit rays
alike.
"""
import ctypes as C

import numpy as np
import pytest

import yune_amd as Y
from yune_amd._lib import lib

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2
YN_ERR_INVALID = -2


def hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            h = C.CDLL(name)
            break
        except OSError:
            continue
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


class DevArray:
    """A device buffer holding a copy of a numpy array's shape and dtype."""

    def __init__(self, h, like: np.ndarray):
        self.h, self.shape, self.dtype = h, like.shape, like.dtype
        self.nbytes = like.nbytes
        self.p = C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), max(self.nbytes, 16)) == 0

    def put(self, a: np.ndarray):
        a = np.ascontiguousarray(a, self.dtype)
        assert self.h.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        assert self.h.hipMemcpy(out.ctypes.data, self.p, self.nbytes, D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.p)


def random_rays(n, seed, lo, hi, shadow_frac=0.3):
    rng = np.random.default_rng(seed)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = rng.uniform(lo, hi, size=(n, 3))
    d = rng.normal(size=(n, 3))
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    shadow = rng.uniform(size=n) < shadow_frac
    rays[:, 3] = np.where(shadow, rng.uniform(0.05, 6.0, size=n), np.inf)
    rays[:, 7] = shadow
    return rays


def trace_noalloc(c, h, sid, rays, n_buf):
    """yn_trace_noalloc on the first len(rays) entries of device buffers
    sized for n_buf rays; returns (tri, light, t) of those rays."""
    n = len(rays)
    d_rays = DevArray(h, np.zeros((n_buf, 8), np.float32))
    d_tri = DevArray(h, np.zeros(n_buf, np.int32))
    d_light = DevArray(h, np.zeros(n_buf, np.int32))
    d_t = DevArray(h, np.zeros(n_buf, np.float32))
    try:
        pad = np.zeros((n_buf, 8), np.float32)
        pad[:n] = rays
        d_rays.put(pad)
        d_tri.put(np.full(n_buf, -7, np.int32))
        r = lib().yn_trace_noalloc(c._p, sid, d_rays.p, n, d_tri.p, d_light.p, d_t.p)
        assert r == 0, lib().yn_last_error(c._p)
        tri, light, t = d_tri.get(), d_light.get(), d_t.get()
        assert (tri[n:] == -7).all(), "wrote past the n rays it was given"
        return tri[:n], light[:n], t[:n]
    finally:
        for b in (d_rays, d_tri, d_light, d_t):
            b.free()


@pytest.mark.parametrize("scene", ["cornell", "ico"])
def test_noalloc_trace_equals_yn_trace(scene):
    if scene == "cornell":
        s = Y.gen_cornell()
        lo, hi = (-0.99, -0.99, -3.99), (0.99, 0.99, -2.01)
    else:
        s = Y.gen_icosphere_room(11, 24, 4)
        lo, hi = (-4.9, -4.9, -4.9), (4.9, 4.9, 4.9)
    h = hip()
    print(f"\nno-alloc trace: scene {scene}, {len(s.tris)} triangles", flush=True)
    with Y.Context(max_paths=1 << 20) as c:
        sid = c.scene(s)
        for k, n in enumerate((200_000, 3_000)):
            rays = random_rays(n, 40 + k, lo, hi)
            want = c.trace(sid, rays)
            got = trace_noalloc(c, h, sid, rays, 200_000)
            for a, b, what in zip(got, want, ("tri", "light", "t")):
                assert np.array_equal(a.view(np.int32), b.view(np.int32)), f"call {k} ({n} rays): {what} differs"
            shadow = rays[:, 7] != 0
            hit = got[0] >= 0
            print(f"  call {k}: {n} rays, {shadow.sum()} any-hit, {(hit & ~shadow).sum()} closest hits", flush=True)
            assert hit[~shadow].any() and hit[shadow].any()


def test_noalloc_refuses_more_rays_than_its_scratch_holds():
    s = Y.gen_cornell()
    h = hip()
    n_max = 262144
    with Y.Context(max_paths=1 << 16) as c:
        sid = c.scene(s)
        buf = DevArray(h, np.zeros(16, np.float32))
        try:
            r = lib().yn_trace_noalloc(c._p, sid, buf.p, n_max + 1, buf.p, buf.p, buf.p)
        finally:
            buf.free()
        assert r == YN_ERR_INVALID
        assert b"at most 262144 rays" in lib().yn_last_error(c._p)
        assert lib().yn_trace_noalloc(c._p, sid, None, 0, None, None, None) == 0
