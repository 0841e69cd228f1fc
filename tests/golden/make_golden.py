"""Writes tests/golden/bvh_golden.json (+ small .npz node arrays).

The expected node arrays come from the reference's own SAH builder
(src/BVH.cpp, src/BVHNodeCPU.cpp, src/TriangleCPU.cpp compiled into
oracle/_ref by `make -C oracle ref`), run on:
  * the procedural 36-triangle Cornell box and icosphere rooms made by
    yune_amd's scene generator, up to the 983,054-triangle benchmark scene
    (the input triangles are recorded by hash, so a generator change shows
    up as an input mismatch, not a BVH one);
  * the reference's Cornell fixtures geometry/cornellbox.obj and
    geometry/cornellbox-teapot.obj, parsed by the independent reader in
    tests/objread.py.

It also writes cornellbox_scene.npz: geometry/cornellbox.obj with its .mtl
(the reference's own input fixture) as the kernel arrays Scene::loadModel
makes of it (triangles, materials, light quad, camera; read with
yn_load_obj, whose parsing tests/test_scene_io.py checks against the
independent reader), so GPU tests can render the reference's scene where
the reference checkout is absent.

Run from the repository root:  python tests/golden/make_golden.py
(`... make_golden.py scenes` writes only cornellbox_scene.npz)
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import yune_amd as Y  # noqa: E402
from oracle import pyoracle as O  # noqa: E402
import objread  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def tri_digest(tris):
    t = np.ascontiguousarray(tris, dtype=Y.TRI_DTYPE)
    return hashlib.sha256(t.tobytes()).hexdigest()


def node_digest(nodes):
    """SHA-256 of the fields the reference defines: box xyz, child_idx,
    vert_len and vert_list[0:vert_len] (the rest is uninitialised memory in
    the reference, BVHNodeCPU.cpp:30-38)."""
    h = hashlib.sha256()
    for x in nodes:
        vl = int(x["vert_len"])
        h.update(np.ascontiguousarray(x["p_min"][:3]).tobytes())
        h.update(np.ascontiguousarray(x["p_max"][:3]).tobytes())
        h.update(np.int32(x["child_idx"]).tobytes())
        h.update(np.int32(vl).tobytes())
        h.update(np.ascontiguousarray(x["vert_list"][:max(vl, 0)]).tobytes())
    return h.hexdigest()


def cases():
    yield "cornell36", Y.gen_cornell().tris, {"generator": "gen_cornell"}
    for seed, n, sub in ((7, 4, 3), (11, 6, 3), (3, 9, 2), (5, 3, 4), (2024, 48, 5)):
        yield f"ico_{seed}_{n}_{sub}", Y.gen_icosphere_room(seed, n, sub).tris, {
            "generator": "gen_icosphere_room", "seed": seed, "n_spheres": n, "subdiv": sub}
    for name in ("cornellbox.obj", "cornellbox-teapot.obj"):
        path = os.path.join(REF, "geometry", name)
        yield name, objread.read_obj(path)[0], {"reference_fixture": "geometry/" + name}


def write_scenes():
    s = Y.load_obj(os.path.join(REF, "geometry", "cornellbox.obj"))
    np.savez_compressed(os.path.join(HERE, "cornellbox_scene.npz"), tris=s.tris, mats=s.mats,
                        light=np.frombuffer(bytes(s.light), np.uint8), cam=np.frombuffer(bytes(s.cam), np.uint8))
    print("cornellbox_scene.npz", len(s.tris), "triangles", len(s.mats), "materials")


def main():
    write_scenes()
    if not O.ref_available():
        sys.exit("oracle/_ref/libref_bvh.so is missing: make -C oracle ref (needs the reference checkout)")
    out = {"note": "expected BVH node arrays from the reference's own BVH.cpp (oracle/_ref), bins=20",
           "cases": {}}
    for name, tris, meta in cases():
        nodes = O.ref_bvh_build(tris, 20)
        entry = dict(meta, n_tri=int(len(tris)), n_nodes=int(len(nodes)), tris_sha256=tri_digest(tris),
                     nodes_sha256=node_digest(nodes))
        if len(nodes) <= 2500:
            fn = name.replace(".obj", "") + "_nodes.npz"
            np.savez_compressed(os.path.join(HERE, fn), nodes=nodes)
            entry["nodes_file"] = fn
        out["cases"][name] = entry
        print(name, entry["n_tri"], entry["n_nodes"])
    with open(os.path.join(HERE, "bvh_golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    if sys.argv[1:] == ["scenes"]:
        write_scenes()
    else:
        main()
