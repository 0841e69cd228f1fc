"""Scene preparation on the host: camera packing, loader defaults and the
procedural scenes the benchmarks use."""
import math
import os

import numpy as np
import pytest

import yune_amd as Y
from oracle import pyoracle as O

REF = "/root/reference/geometry"


def test_camera_packing_matches_oracle_bitwise():
    """yn_camera and the oracle's Camera::setViewMatrix / setBuffer
    restatement (Camera.cpp:49, 66-80, 105-117) give the same bytes."""
    rng = np.random.default_rng(1)
    for _ in range(50):
        side, up, look, eye = rng.normal(size=(4, 3)).astype(np.float32)
        fov = float(rng.uniform(10, 120))
        assert bytes(Y.camera(side, up, look, eye, fov)) == bytes(O.camera(side, up, look, eye, fov))


def test_default_camera():
    c = Y.camera((1, 0, 0), (0, 1, 0), (0, 0, -1), (0, 0, 0), 60.0)
    rows = np.array([list(c.r1), list(c.r2), list(c.r3), list(c.r4)], np.float32)
    assert np.array_equal(rows, np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], np.float32))
    assert c.view_plane_dist == np.float32(1 / math.tan(60 * 3.14 / 360))


def check_scene(s):
    assert len(s.tris) > 0 and len(s.mats) > 0
    assert (s.tris["v1"][:, 3] == 1).all() and (s.tris["v2"][:, 3] == 1).all() and (s.tris["v3"][:, 3] == 1).all()
    assert (s.tris["vn1"][:, 3] == 0).all()
    assert (s.tris["matID"] >= 0).all() and (s.tris["matID"] < len(s.mats)).all()
    lq = s.light
    assert lq.pos[3] == 1 and lq.normal[3] == 0 and lq.edge_l[3] == 0 and lq.edge_w[3] == 0
    assert abs(math.sqrt(sum(lq.normal[i] ** 2 for i in range(3))) - 1) < 1e-6


def test_cornell():
    s = Y.gen_cornell()
    check_scene(s)
    assert len(s.tris) == 36
    # the light quad faces into the room and lies under the emissive panel
    assert tuple(s.light.normal)[:3] == (0, -1, 0)
    panel = s.tris[[s.mats[m]["ke"][:3].max() > 0 for m in s.tris["matID"]]]
    assert len(panel) == 2 and panel["v1"][:, 1].min() > s.light.pos[1]


@pytest.mark.parametrize("seed,n,sub", [(7, 4, 3), (2024, 48, 5)])
def test_icosphere_room_counts(seed, n, sub):
    s = Y.gen_icosphere_room(seed, n, sub)
    check_scene(s)
    assert len(s.tris) == n * 20 * 4 ** sub + 14         # spheres + 12 wall triangles + 2 light panel
    assert bytes(s.tris) == bytes(Y.gen_icosphere_room(seed, n, sub).tris)   # deterministic in the seed
    if n * 20 * 4 ** sub < 100000:
        assert bytes(s.tris) != bytes(Y.gen_icosphere_room(seed + 1, n, sub).tris)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference fixtures not present")
def test_obj_defaults_are_the_reference_kernel_constants():
    s = Y.load_obj(os.path.join(REF, "cornellbox.obj"))
    check_scene(s)
    lq = s.light
    assert tuple(lq.pos) == tuple(np.float32([-0.1979, 0.92, -3.1972, 1]))     # udpt.cl:96-104
    assert tuple(lq.ke)[:3] == (16, 16, 16) and tuple(lq.edge_l)[:3] == (np.float32(0.4), 0, 0)
    assert bytes(s.cam) == bytes(Y.camera((1, 0, 0), (0, 1, 0), (0, 0, -1), (0, 0, 0), 60.0))


def test_cornellbox_fixture_is_the_reference_scene():
    """tests/golden/cornellbox_scene.npz (what GPU tests render where the
    reference checkout is absent) holds geometry/cornellbox.obj as loaded:
    the same triangle bytes the golden BVH case was built from."""
    import hashlib
    import json
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    d = np.load(os.path.join(here, "cornellbox_scene.npz"))
    case = json.load(open(os.path.join(here, "bvh_golden.json")))["cases"]["cornellbox.obj"]
    assert hashlib.sha256(np.ascontiguousarray(d["tris"]).tobytes()).hexdigest() == case["tris_sha256"]
    if os.path.isdir(REF):
        s = Y.load_obj(os.path.join(REF, "cornellbox.obj"))
        assert bytes(s.tris) == d["tris"].tobytes() and bytes(s.mats) == d["mats"].tobytes()
        assert bytes(s.light) == d["light"].tobytes() and bytes(s.cam) == d["cam"].tobytes()
