"""yune_amd — MI355X path-tracing hot path of Yune behind a C ABI.

Python mirror of the reference's host side for tests and benchmarks:

  build_bvh          BVH::createBVH            (src/BVH.cpp:56-173)
  load_obj           Scene::loadModel          (src/Scene.cpp:133-383)
  camera             Camera::setViewMatrix + setBuffer (src/Camera.cpp:66-117)
  Context            CLManager                 (src/CLManager.cpp)
  Context.scene      setupVertex/Mat/BVHBuffer (src/CLManager.cpp:420-484)
  Context.job / Job  RendererCore frame loop   (src/RendererCore.cpp:248-579)
  Context.pathtracer the 14-argument kernel    (kernels/legacy/udpt.cl:158-160)
  Context.trace      traceRay                  (kernels/legacy/udpt.cl:240-286)
  Context.submit / serve  many jobs per wavefront batch (no reference counterpart)

All work is done by libyune_hip.so; nothing here computes a result itself.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._lib import (MAT_DTYPE, NODE_DTYPE, TONEMAP_HABLE, TONEMAP_REINHARD, TRI_DTYPE, YN_MODE_BDPT, YN_MODE_PT, Cam,
                   CamCtl, Config, HostScene,
                   JobStatus, TuneResult,
                   LightQuad, RenderParams, SceneInfo, Stats, YuneError, check, lib, ptr)

__all__ = ["TRI_DTYPE", "MAT_DTYPE", "NODE_DTYPE", "YN_MODE_PT", "YN_MODE_BDPT", "TONEMAP_REINHARD", "TONEMAP_HABLE",
           "YuneError", "Scene",
           "build_bvh", "load_obj", "gen_cornell", "gen_icosphere_room", "camera", "params_reference",
           "Context", "Job", "light_quad"]


def params_reference(mode: int = YN_MODE_PT, **over) -> RenderParams:
    """udpt.cl / bdpt.cl defaults, with keyword overrides."""
    p = RenderParams()
    lib().yn_params_reference(C.byref(p), mode)
    for k, v in over.items():
        if k == "background":
            for i, x in enumerate(v):
                p.background[i] = x
        else:
            if not hasattr(p, k):
                raise AttributeError(k)
            setattr(p, k, v)
    return p


def build_gpu_bvh(tris: np.ndarray, leaf_max: int = 2) -> np.ndarray:
    """The library's own SAH tree (what a context with bvh_build=1 lays out),
    in the reference's node format."""
    tris = np.ascontiguousarray(tris, dtype=TRI_DTYPE)
    out = C.c_void_p()
    n = C.c_int32()
    check(lib().yn_bvh_build_gpu(ptr(tris), len(tris), leaf_max, C.byref(out), C.byref(n)))
    try:
        buf = (C.c_char * (n.value * NODE_DTYPE.itemsize)).from_address(out.value)
        return np.frombuffer(bytes(buf), dtype=NODE_DTYPE).copy()
    finally:
        lib().yn_free(out)


def build_bvh(tris: np.ndarray, bins: int = 20) -> np.ndarray:
    tris = np.ascontiguousarray(tris, dtype=TRI_DTYPE)
    out = C.c_void_p()
    n = C.c_int32()
    check(lib().yn_bvh_build(ptr(tris), len(tris), bins, C.byref(out), C.byref(n)))
    try:
        buf = (C.c_char * (n.value * NODE_DTYPE.itemsize)).from_address(out.value)
        return np.frombuffer(bytes(buf), dtype=NODE_DTYPE).copy()
    finally:
        lib().yn_free(out)


@dataclass
class Scene:
    """Host-side scene: the triangle / material arrays the reference uploads,
    the light quad of udpt.cl:97-106, and a camera."""
    tris: np.ndarray
    mats: np.ndarray
    light: LightQuad
    cam: Cam
    nodes: np.ndarray | None = field(default=None)

    def bvh(self, bins: int = 20) -> np.ndarray:
        if self.nodes is None:
            self.nodes = build_bvh(self.tris, bins)
        return self.nodes


def _from_host(h: HostScene) -> Scene:
    try:
        tb = (C.c_char * (h.n_tri * TRI_DTYPE.itemsize)).from_address(h.tris)
        mb = (C.c_char * (h.n_mat * MAT_DTYPE.itemsize)).from_address(h.mats)
        tris = np.frombuffer(bytes(tb), dtype=TRI_DTYPE).copy()
        mats = np.frombuffer(bytes(mb), dtype=MAT_DTYPE).copy()
        light = LightQuad.from_buffer_copy(h.light)
        cam = Cam.from_buffer_copy(h.cam)
    finally:
        lib().yn_host_scene_free(C.byref(h))
    return Scene(tris, mats, light, cam)


def load_obj(path: str) -> Scene:
    h = HostScene()
    check(lib().yn_load_obj(path.encode(), C.byref(h)))
    return _from_host(h)


def gen_cornell() -> Scene:
    h = HostScene()
    check(lib().yn_gen_cornell(C.byref(h)))
    return _from_host(h)


def gen_icosphere_room(seed: int, n_spheres: int = 48, subdiv: int = 5) -> Scene:
    h = HostScene()
    check(lib().yn_gen_icosphere_room(seed, n_spheres, subdiv, C.byref(h)))
    return _from_host(h)


def camera(side, up, look_at, eye, yfov: float = 60.0) -> Cam:
    arrs = [np.ascontiguousarray(v, dtype=np.float32) for v in (side, up, look_at, eye)]
    c = Cam()
    fp = C.POINTER(C.c_float)
    check(lib().yn_camera(*[a.ctypes.data_as(fp) for a in arrs], float(yfov), C.byref(c)))
    return c


class CameraControl:
    """Camera (src/Camera.cpp): move / pitch / yaw controls and setBuffer."""

    def __init__(self, y_fov: float = 60.0, rotation_speed: float = 0.25, move_speed: float = 0.1):
        self.s = CamCtl()
        check(lib().yn_camctl_init(y_fov, rotation_speed, move_speed, C.byref(self.s)))

    def set_view(self, side, up, look_at, eye):
        arrs = [np.ascontiguousarray(v, dtype=np.float32) for v in (side, up, look_at, eye)]
        fp = C.POINTER(C.c_float)
        check(lib().yn_camctl_set_view(C.byref(self.s), *[a.ctypes.data_as(fp) for a in arrs]))

    def orient(self, direction=(0.0, 0.0, 0.0), pitch: float = 0.0, yaw: float = 0.0):
        d = np.ascontiguousarray(direction, dtype=np.float32)
        check(lib().yn_camctl_orient(C.byref(self.s), d.ctypes.data_as(C.POINTER(C.c_float)), pitch, yaw))

    def cam(self) -> Cam:
        c = Cam()
        check(lib().yn_camctl_pack(C.byref(self.s), C.byref(c)))
        return c

    def basis(self):
        g = lambda v: np.array(list(v), np.float32)  # noqa: E731
        return g(self.s.side), g(self.s.up), g(self.s.look_at), g(self.s.eye)


def light_quad(pos, normal, ke, edge_l, edge_w) -> LightQuad:
    q = LightQuad()
    for name, v, w in (("pos", pos, 1.0), ("normal", normal, 0.0), ("ke", ke, 0.0), ("edge_l", edge_l, 0.0),
                       ("edge_w", edge_w, 0.0)):
        arr = getattr(q, name)
        for i in range(3):
            arr[i] = float(v[i])
        arr[3] = w
    return q


class Context:
    """A device context (CLManager).  Raises YuneError when no gfx950 GPU is
    present; there is no CPU fallback."""

    def __init__(self, device: int = 0, max_paths: int = 0, bvh_width: int = 0, trace_block: int = 0,
                 trace_chunk: int = 0, trace_grid: int = 0, mem_budget: int = 0, hoist_frac: float = 0.0,
                 leaf_max: int = 0, lds_stack: int = 0, shade_waves: int = 0, batch_factor: int = 0,
                 trace_overlap: int = 0, graph: int = 0, trace_xcd: int = 0, trace_refill: int = 0,
                 bvh_build: int = 0, serve_mode: int = 0):
        cfg = Config(device, max_paths, mem_budget, bvh_width, trace_block, trace_chunk, trace_grid, hoist_frac,
                     leaf_max, lds_stack, shade_waves, batch_factor, trace_overlap, graph, trace_xcd, trace_refill,
                     bvh_build, serve_mode)
        self._p = C.c_void_p()
        check(lib().yn_create(C.byref(cfg), C.byref(self._p)))
        self.cfg = cfg

    def close(self):
        if self._p:
            lib().yn_destroy(self._p)
            self._p = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, code):
        check(code, self._p)

    def scene(self, s: Scene, nodes: np.ndarray | None = None, brute: bool = False) -> int:
        """Upload a scene; its reference BVH is built unless `nodes` is
        given.  brute=True uploads no BVH (kernel arg bvh_size == 0)."""
        if brute:
            nodes = np.zeros(0, dtype=NODE_DTYPE)
        elif nodes is None:
            nodes = s.bvh()
        nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
        tris = np.ascontiguousarray(s.tris, dtype=TRI_DTYPE)
        mats = np.ascontiguousarray(s.mats, dtype=MAT_DTYPE)
        sid = C.c_int32()
        self._ck(lib().yn_scene_create(self._p, ptr(tris), len(tris), ptr(mats), len(mats), ptr(nodes),
                                       len(nodes), C.byref(s.light), C.byref(sid)))
        return sid.value

    def scene_info(self, sid: int) -> SceneInfo:
        info = SceneInfo()
        self._ck(lib().yn_scene_info_get(self._p, sid, C.byref(info)))
        return info

    def release_scene(self, sid: int):
        self._ck(lib().yn_scene_release(self._p, sid))

    def job(self, sid: int, cam: Cam, width: int, height: int, params: RenderParams, base_seed: int = 1) -> "Job":
        jid = C.c_int32()
        self._ck(lib().yn_job_create(self._p, sid, C.byref(cam), width, height, C.byref(params), base_seed,
                                     C.byref(jid)))
        return Job(self, jid.value, width, height)

    def pathtracer(self, out_img, in_img, cam: Cam, sid: int, width: int, height: int, params: RenderParams,
                   gi_check: int, reset: int, rand: int, block: int = 0, block_x: int = 1, block_y: int = 1):
        """One frame over tile `block` of a block_x x block_y grid on device
        images (torch float32 tensors of H x W x 4 on the GPU), with the
        meaning of the reference kernel's arguments 0-13."""
        import torch
        for t in (out_img,) + ((in_img,) if in_img is not None else ()):
            if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and t.numel() == width * height * 4):
                raise ValueError("images must be contiguous float32 CUDA tensors of height x width x 4")
        torch.cuda.synchronize()            # the library works on its own stream
        self._ck(lib().yn_pathtracer(self._p, C.c_void_p(out_img.data_ptr()),
                                     C.c_void_p(in_img.data_ptr()) if in_img is not None else None,
                                     C.byref(cam), sid, width, height, C.byref(params), gi_check, reset, rand,
                                     block, block_x, block_y))

    def tonemap(self, out_img, in_img, width: int, height: int, op: int = 0):
        """tonemap.cl on device images (H x W x 4 float32 CUDA tensors):
        op TONEMAP_REINHARD (as shipped) or TONEMAP_HABLE."""
        import torch
        for t in (out_img, in_img):
            if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and t.numel() == width * height * 4):
                raise ValueError("images must be contiguous float32 CUDA tensors of height x width x 4")
        torch.cuda.synchronize()
        self._ck(lib().yn_tonemap_op(self._p, C.c_void_p(in_img.data_ptr()), C.c_void_p(out_img.data_ptr()),
                                     width, height, op))

    def submit(self, sid: int, cam: Cam, width: int, height: int, params: RenderParams, frames: int,
               base_seed: int = 1) -> "Job":
        """Queue a job of `frames` frames for the scheduler (yn_submit)."""
        jid = C.c_int32()
        self._ck(lib().yn_submit(self._p, sid, C.byref(cam), width, height, C.byref(params), base_seed, frames,
                                 C.byref(jid)))
        return Job(self, jid.value, width, height)

    def serve(self, max_rounds: int = 1) -> int:
        """Run up to max_rounds scheduler steps (serve_mode 0: groups of four
        bounce iterations of the continuous loop; 1: rounds); returns the
        jobs still pending."""
        pend = C.c_int32()
        self._ck(lib().yn_serve(self._p, max_rounds, C.byref(pend)))
        return pend.value

    def serve_stats(self) -> Stats:
        st = Stats()
        self._ck(lib().yn_serve_stats(self._p, C.byref(st)))
        return st

    def autotune(self, sid: int, cam: Cam, width: int, height: int, params: RenderParams, frames: int = 2) -> dict:
        """Choose BVH width, launch shape, LDS stack depth, shade occupancy and
        batch size for this view (yn_autotune); the context keeps them."""
        r = TuneResult()
        self._ck(lib().yn_autotune(self._p, sid, C.byref(cam), width, height, C.byref(params), frames, C.byref(r)))
        return {k: getattr(r, k) for k, _ in TuneResult._fields_}

    def config(self) -> dict:
        cfg = Config()
        bp = C.c_int64()
        self._ck(lib().yn_config_get(self._p, C.byref(cfg), C.byref(bp)))
        d = {k: getattr(cfg, k) for k, _ in Config._fields_}
        d["batch_paths"] = bp.value
        return d

    def clock_ms(self) -> float:
        return lib().yn_clock_ms(self._p)

    def trace(self, sid: int, rays: np.ndarray):
        """rays: (n, 8) float32 (o.xyz, length, d.xyz, is_shadow)."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = len(rays)
        tri = np.empty(n, np.int32)
        light = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        self._ck(lib().yn_trace(self._p, sid, ptr(rays), n, ptr(tri), ptr(light), ptr(t)))
        return tri, light, t

    def device_info(self):
        name = C.create_string_buffer(64)
        ncu = C.c_int32()
        hbm = C.c_size_t()
        self._ck(lib().yn_device_info(self._p, name, 64, C.byref(ncu), C.byref(hbm)))
        return name.value.decode(), ncu.value, hbm.value


class Job:
    """A progressive render (RendererCore's frame loop, headless)."""

    def __init__(self, ctx: Context, jid: int, width: int, height: int):
        self.ctx, self.id, self.width, self.height = ctx, jid, width, height

    def render(self, frames: int):
        self.ctx._ck(lib().yn_job_render(self.ctx._p, self.id, frames))

    def read(self) -> np.ndarray:
        img = np.empty((self.height, self.width, 4), np.float32)
        self.ctx._ck(lib().yn_job_read(self.ctx._p, self.id, ptr(img)))
        return img

    def read_tonemapped(self) -> np.ndarray:
        img = np.empty((self.height, self.width, 4), np.float32)
        self.ctx._ck(lib().yn_job_read_tonemapped(self.ctx._p, self.id, ptr(img)))
        return img

    def set_tonemap(self, op: int):
        """TONEMAP_REINHARD (the shipped tonemap.cl) or TONEMAP_HABLE, for
        read_tonemapped and .png saves."""
        self.ctx._ck(lib().yn_job_set_tonemap(self.ctx._p, self.id, op))

    def save(self, path: str):
        self.ctx._ck(lib().yn_job_save(self.ctx._p, self.id, path.encode()))

    def stats(self) -> Stats:
        st = Stats()
        self.ctx._ck(lib().yn_job_stats(self.ctx._p, self.id, C.byref(st)))
        return st

    def status(self) -> dict:
        st = JobStatus()
        self.ctx._ck(lib().yn_job_status_get(self.ctx._p, self.id, C.byref(st)))
        return {k: getattr(st, k) for k, _ in JobStatus._fields_}

    def set_counting(self, on: bool):
        self.ctx._ck(lib().yn_job_set_counting(self.ctx._p, self.id, int(on)))

    def capture_next(self):
        """Arrays filled by the next render() call (its first frame)."""
        n = self.width * self.height
        cap = {"prim_tri": np.empty(n, np.int32), "prim_light": np.empty(n, np.int32),
               "prim_t": np.empty(n, np.float32), "shadow_ray": np.empty((n, 8), np.float32),
               "shadow_occ": np.empty(n, np.int32)}
        self.ctx._ck(lib().yn_job_capture(self.ctx._p, self.id, ptr(cap["prim_tri"]), ptr(cap["prim_light"]),
                                          ptr(cap["prim_t"]), ptr(cap["shadow_ray"]), ptr(cap["shadow_occ"])))
        self._cap = cap
        return cap

    def release(self):
        self.ctx._ck(lib().yn_job_release(self.ctx._p, self.id))
