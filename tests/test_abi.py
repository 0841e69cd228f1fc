"""The C-ABI boundary, checked without a GPU.

* libyune_hip.so loads and exports every function include/yune_hip.h
  declares, and the Python binding declares each of them;
* the ctypes / numpy mirrors have the byte layout gcc gives the C structs
  (a small C program compiled here prints sizeof/offsetof);
* host-side entry points validate their arguments, and yn_create reports
  YN_ERR_NO_DEVICE (not a crash, not a CPU fallback) where no GPU exists.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest
import torch

import yune_amd as Y
from yune_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDRS = [os.path.join(ROOT, "include", h) for h in ("yune_types.h", "yune_hip.h")]


def declared_functions():
    src = open(HDRS[1]).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(yn_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported_and_bound():
    names = declared_functions()
    assert len(names) >= 25
    assert sorted(L.SIGNATURES) == names
    lib = L.lib()
    for n in names:
        assert getattr(lib, n) is not None
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (yn_[a-z0-9_]+)$", out, flags=re.M))
    assert set(names) <= exported
    # nothing else of ours leaks into the dynamic symbol table
    assert exported == set(names)


LAYOUT_C = r"""
#include <stdio.h>
#include <stddef.h>
#include "yune_hip.h"
#define S(T) printf(#T " sizeof %zu\n", sizeof(T));
#define O(T, f) printf(#T " " #f " %zu\n", offsetof(T, f));
int main(void) {
  S(yn_cam) O(yn_cam, view_plane_dist)
  S(yn_triangle) O(yn_triangle, vn1) O(yn_triangle, matID)
  S(yn_bvh_node) O(yn_bvh_node, vert_list) O(yn_bvh_node, child_idx) O(yn_bvh_node, vert_len)
  S(yn_material) O(yn_material, n) O(yn_material, px) O(yn_material, alpha_x) O(yn_material, is_specular)
    O(yn_material, is_transmissive)
  S(yn_light_quad) O(yn_light_quad, ke) O(yn_light_quad, edge_w)
  S(yn_render_params) O(yn_render_params, fix_mis) O(yn_render_params, background)
  S(yn_config) O(yn_config, mem_budget) O(yn_config, bvh_width) O(yn_config, hoist_frac) O(yn_config, leaf_max) \
    O(yn_config, lds_stack) O(yn_config, shade_waves) O(yn_config, batch_factor) O(yn_config, trace_overlap) O(yn_config, graph) O(yn_config, trace_xcd) O(yn_config, trace_refill) O(yn_config, bvh_build)
  S(yn_host_scene) O(yn_host_scene, n_tri) O(yn_host_scene, mats) O(yn_host_scene, n_mat)
    O(yn_host_scene, light) O(yn_host_scene, cam)
  S(yn_scene_info) O(yn_scene_info, n_hoist) O(yn_scene_info, node_bytes) O(yn_scene_info, shade_bytes)
  S(yn_stats) O(yn_stats, iterations) O(yn_stats, ms_total) O(yn_stats, ms_other)
  S(yn_tune_result) O(yn_tune_result, batch_paths) O(yn_tune_result, graph) O(yn_tune_result, trace_xcd) O(yn_tune_result, trace_refill) O(yn_tune_result, mrays_per_s)
  S(yn_job_status) S(yn_camctl)
  return 0;
}
"""


def test_struct_layout_matches_gcc(tmp_path):
    (tmp_path / "lay.c").write_text(LAYOUT_C)
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-o", str(tmp_path / "lay"),
                    str(tmp_path / "lay.c")], check=True)
    got = {}
    for line in subprocess.run([str(tmp_path / "lay")], capture_output=True, text=True, check=True).stdout.split("\n"):
        if line:
            *k, v = line.split()
            got[tuple(k)] = int(v)

    def off(struct, field):
        return getattr(struct, field).offset

    np_types = {"yn_triangle": Y.TRI_DTYPE, "yn_bvh_node": Y.NODE_DTYPE, "yn_material": Y.MAT_DTYPE}
    ct_types = {"yn_cam": L.Cam, "yn_light_quad": L.LightQuad, "yn_render_params": L.RenderParams,
                "yn_config": L.Config, "yn_host_scene": L.HostScene, "yn_scene_info": L.SceneInfo,
                "yn_stats": L.Stats, "yn_tune_result": L.TuneResult, "yn_job_status": L.JobStatus,
                "yn_camctl": L.CamCtl}
    for key, v in got.items():
        name = key[0]
        if name in np_types:
            dt = np_types[name]
            exp = dt.itemsize if key[1] == "sizeof" else dt.fields[key[1]][1]
        else:
            ct = ct_types[name]
            exp = C.sizeof(ct) if key[1] == "sizeof" else off(ct, key[1])
        assert exp == v, (key, exp, v)


def test_params_reference_defaults():
    p = Y.params_reference()
    assert (p.mode, p.gi, p.mis, p.max_bounces, p.rr_threshold) == (0, 1, 0, 100000, 6)
    assert (p.bdpt_bounces, p.oren_nayar, p.fix_mis) == (20, 0, 0)
    assert list(p.background) == [np.float32(0.4)] * 3 + [1.0]
    b = Y.params_reference(Y.YN_MODE_BDPT)
    assert (b.mode, b.rr_threshold) == (1, 4)


def test_no_device_is_an_error_not_a_fallback():
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert L.lib().yn_create(None, C.byref(h)) == -6
    assert not h.value
    assert b"device" in L.lib().yn_last_error(None)
    with pytest.raises(Y.YuneError) as e:
        Y.Context()
    assert e.value.code == -6


def test_config_validation_precedes_device_probe():
    cfg = L.Config(bvh_width=3)
    h = C.c_void_p()
    assert L.lib().yn_create(C.byref(cfg), C.byref(h)) == -2
    assert b"bvh_width" in L.lib().yn_last_error(None)
    cfg = L.Config(trace_block=96)
    assert L.lib().yn_create(C.byref(cfg), C.byref(h)) == -2
    assert L.lib().yn_create(C.byref(cfg), None) == -2


def test_null_context_is_rejected():
    lib = L.lib()
    st = L.Stats()
    assert lib.yn_job_stats(None, 0, C.byref(st)) != 0
    assert lib.yn_scene_release(None, 0) != 0
    assert lib.yn_tonemap(None, None, None, 4, 4) == -2
    n = C.c_int32()
    assert lib.yn_device_info(None, None, 0, C.byref(n), None) == -2


def test_host_entry_points_validate_arguments():
    lib = L.lib()
    nodes, n = C.c_void_p(), C.c_int32()
    assert lib.yn_bvh_build(None, -1, 20, C.byref(nodes), C.byref(n)) == -2
    assert lib.yn_bvh_build(None, 5, 20, C.byref(nodes), C.byref(n)) == -2
    assert lib.yn_bvh_build(None, 0, 20, None, C.byref(n)) == -2
    assert lib.yn_bvh_build_gpu(None, 5, 2, C.byref(nodes), C.byref(n)) == -2
    assert lib.yn_bvh_build_gpu(None, 0, 11, C.byref(nodes), C.byref(n)) == -2
    assert lib.yn_gen_icosphere_room(1, -1, 3, C.byref(L.HostScene())) == -2
    assert lib.yn_gen_icosphere_room(1, 2, 9, C.byref(L.HostScene())) == -2
    assert lib.yn_load_obj(None, C.byref(L.HostScene())) == -2
    assert lib.yn_camera(None, None, None, None, 60.0, None) == -2
    with pytest.raises(Y.YuneError):
        Y.load_obj("/nonexistent/scene.obj")


def test_empty_scene_bvh():
    """No triangles: one root marked empty (child_idx -2, vert_len -1), as
    the reference's builder leaves it (checked against it in test_bvh)."""
    nodes = Y.build_bvh(np.zeros(0, Y.TRI_DTYPE))
    assert len(nodes) == 1 and nodes[0]["child_idx"] == -2 and nodes[0]["vert_len"] == -1


def test_degenerate_input_is_refused_not_looped():
    """Eleven triangles with one centroid: the reference's builder loops
    forever (every split leaves all of them on one side); ours stops at its
    node limit and reports YN_ERR_INVALID."""
    t = np.zeros(11, Y.TRI_DTYPE)
    for i in range(11):
        t[i]["v1"] = (0, 0, 0, 1)
        t[i]["v2"] = (1, 0, 0, 1)
        t[i]["v3"] = (0, 1, 0, 1)
    with pytest.raises(Y.YuneError) as e:
        Y.build_bvh(t)
    assert e.value.code == -2


def test_product_path_has_no_oracle_dependency():
    """The product package never imports the oracle (it is test
    infrastructure) and links only HIP runtime libraries."""
    for dirpath, _, files in os.walk(os.path.join(ROOT, "yune_amd")):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                txt = open(os.path.join(dirpath, f), errors="replace").read()
                assert "oracle" not in txt.replace("the oracle", ""), f
    out = subprocess.run(["ldd", L.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out and "libamdhip64" in out
