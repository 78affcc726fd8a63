"""glTF ingestion (gpurt.gltf) on the CPU: the reference's data-only scenes against facts
an independent reader recorded (tests/golden/gltf_facts.json, made by
tests/golden/make_gltf_golden.py), and the fan/strip/index-type/attribute rules of
scene.cpp:48-315 on synthetic files."""
import base64
import json
import os

import numpy as np
import pytest

from gpurt import gltf

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MEDIA = "/root/reference/media"


@pytest.mark.parametrize("name,rel", [("cube", "cube.gltf"), ("mis_test", "mis_test/mis_test.gltf"),
                                      ("cbox", "cbox/cbox.gltf")])
def test_reference_scenes_match_independent_reader(name, rel):
    path = os.path.join(MEDIA, rel)
    if not os.path.exists(path):
        pytest.skip("reference media not mounted")
    facts = json.load(open(os.path.join(GOLDEN, "gltf_facts.json")))[name]
    objs = gltf.load(path)["objects"]
    assert len(objs) == len(facts)
    fx = np.load(os.path.join(GOLDEN, f"gltf_{name}.npz"))
    for k, (o, f) in enumerate(zip(objs, facts)):
        tri = o["indices"].astype(np.int64)
        assert o["verts"].shape == (f["n_verts"], 12) and tri.shape == (f["n_tris"], 3)
        assert int(tri.sum()) == f["index_sum"]
        assert int((tri * np.arange(1, tri.size + 1).reshape(tri.shape)).sum()) == f["index_checksum"]
        assert np.array_equal(o["verts"][:, :3].min(0), np.float32(f["pos_min"]))
        assert np.array_equal(o["verts"][:, :3].max(0), np.float32(f["pos_max"]))
        assert np.array_equal(o["verts"][:, :3], fx[f"obj{k}_pos"]) and np.array_equal(o["indices"], fx[f"obj{k}_tris"])
        assert np.array_equal(o["transform"][:3].astype(np.float32).reshape(12), fx["xf"][k])
    shading = os.path.join(GOLDEN, f"gltf_{name}_shading.npz")
    if os.path.exists(shading):  # normals and material factors (the path tracer's inputs)
        sh = np.load(shading)
        for k, o in enumerate(objs):
            assert np.array_equal(o["verts"][:, 4:7], sh[f"obj{k}_nrm"])
            m = o["material"]
            assert np.array_equal(m["albedo"], sh["albedo"][k]) and np.array_equal(m["emissive"], sh["emissive"][k])
            assert np.array_equal(m["metal_rough"], sh["metal_rough"][k])


def _synthetic(tmp_path, prims, positions, extra_attrs=None):
    """A one-node glTF with a data: buffer holding `positions` (float32) and each
    primitive's index array (given as (numpy array, componentType, mode))."""
    blob = bytearray()
    views, accessors = [], []

    def add(arr, comp, typ, target):
        nonlocal blob
        while len(blob) % 4:
            blob += b"\0"
        views.append({"buffer": 0, "byteOffset": len(blob), "byteLength": arr.nbytes, "target": target})
        blob += arr.tobytes()
        accessors.append({"bufferView": len(views) - 1, "componentType": comp, "count": int(arr.shape[0]),
                          "type": typ})
        return len(accessors) - 1

    pos = add(np.asarray(positions, np.float32), 5126, "VEC3", 34962)
    attrs = {"POSITION": pos}
    for key, (arr, typ) in (extra_attrs or {}).items():
        attrs[key] = add(np.asarray(arr, np.float32), 5126, typ, 34962)
    primitives = []
    for idx, comp, mode in prims:
        a = add(np.asarray(idx), comp, "SCALAR", 34963)
        primitives.append({"attributes": attrs, "indices": a, "mode": mode})
    doc = {"asset": {"version": "2.0"}, "scenes": [{"nodes": [0]}],
           "nodes": [{"mesh": 0, "matrix": [2, 0, 0, 0, 0, 2, 0, 0, 0, 0, 2, 0, 1, 2, 3, 1]}],
           "meshes": [{"primitives": primitives}], "accessors": accessors, "bufferViews": views,
           "buffers": [{"byteLength": len(blob),
                        "uri": "data:application/octet-stream;base64," + base64.b64encode(bytes(blob)).decode()}]}
    p = tmp_path / "s.gltf"
    p.write_text(json.dumps(doc))
    return str(p)


def test_fans_strips_index_types_and_ignored_modes(tmp_path):
    P = np.arange(18, dtype=np.float32).reshape(6, 3)
    path = _synthetic(tmp_path, [
        (np.array([0, 1, 2, 3, 4, 5], np.uint8), 5121, 6),    # fan, unsigned byte
        (np.array([0, 1, 2, 3, 4, 5], np.uint16), 5123, 5),   # strip, unsigned short
        (np.array([5, 4, 3, 2, 1, 0], np.uint32), 5125, 4),   # list, unsigned int
        (np.array([0, 1], np.uint16), 5123, 1),               # lines: ignored
    ], P)
    objs = gltf.load(path)["objects"]
    assert len(objs) == 3
    assert objs[0]["indices"].tolist() == [[0, 1, 2], [0, 2, 3], [0, 3, 4], [0, 4, 5]]
    # the reference does not alternate the winding of strip triangles (scene.cpp:128-136)
    assert objs[1]["indices"].tolist() == [[0, 1, 2], [1, 2, 3], [2, 3, 4], [3, 4, 5]]
    assert objs[2]["indices"].tolist() == [[5, 4, 3], [2, 1, 0]]
    T = objs[0]["transform"]
    assert np.array_equal(T, [[2, 0, 0, 1], [0, 2, 0, 2], [0, 0, 2, 3], [0, 0, 0, 1]])
    assert np.array_equal(gltf.world_vertices(objs[0]), (P * 2 + [1, 2, 3]).astype(np.float32))


def test_vertex_packing_is_mesh_vertex(tmp_path):
    """{pos.xyz, u}, {normal.xyz, v}, {tangent} per vertex (scene.cpp:305-311)."""
    P = np.arange(9, dtype=np.float32).reshape(3, 3)
    N = np.arange(9, dtype=np.float32).reshape(3, 3) + 100
    Tn = np.arange(12, dtype=np.float32).reshape(3, 4) + 200
    UV = np.arange(6, dtype=np.float32).reshape(3, 2) + 300
    path = _synthetic(tmp_path, [(np.array([0, 1, 2], np.uint16), 5123, 4)], P,
                      {"NORMAL": (N, "VEC3"), "TANGENT": (Tn, "VEC4"), "TEXCOORD_0": (UV, "VEC2")})
    v = gltf.load(path)["objects"][0]["verts"]
    assert v.shape == (3, 12) and v.dtype == np.float32
    assert np.array_equal(v[:, 0:3], P) and np.array_equal(v[:, 3], UV[:, 0])
    assert np.array_equal(v[:, 4:7], N) and np.array_equal(v[:, 7], UV[:, 1])
    assert np.array_equal(v[:, 8:12], Tn)
    # absent attributes are zero
    path2 = _synthetic(tmp_path, [(np.array([0, 1, 2], np.uint16), 5123, 4)], P)
    assert not gltf.load(path2)["objects"][0]["verts"][:, 3:].any()
