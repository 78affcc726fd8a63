"""Writes the glTF golden fixtures (run in a container where the reference is mounted):

    python tests/golden/make_gltf_golden.py /root/reference/media

For media/cube.gltf, media/mis_test/mis_test.gltf and media/cbox/cbox.gltf (the
reference's data-only scenes; it has no tests) it records

* gltf_facts.json: per triangle primitive reached by the scene graph, the facts an
  independent reader computes with plain struct unpacking (not the library's loader):
  vertex and triangle counts, the sum and a position-weighted checksum of the
  indices, and the position bounds;
* gltf_<name>.npz: the object-space positions and index triples of every object, its
  4x4 world matrix, the scene flattened to world space (FP64 transform, rounded to
  FP32), seeded rays and closest-point queries, and the oracle's FP32 brute-force
  answers on the flattened scene (closest hit, any hit, closest point) and on the
  instanced scene (TLAS semantics).

The expected answers come from oracle/ (FP32 restatement pinned by tests/test_oracle.py).
"""
import json
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

FMT = {5120: "b", 5121: "B", 5122: "h", 5123: "H", 5124: "i", 5125: "I", 5126: "f"}
NC = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4}


def read_accessor(doc, bufs, i):
    a = doc["accessors"][i]
    v = doc["bufferViews"][a["bufferView"]]
    fmt = FMT[a["componentType"]]
    size = struct.calcsize("<" + fmt)
    nc = NC[a["type"]]
    stride = v.get("byteStride", 0) or size * nc
    base = v.get("byteOffset", 0) + a.get("byteOffset", 0)
    data = bufs[v["buffer"]]
    out = []
    for k in range(a["count"]):
        out.append(struct.unpack_from("<" + fmt * nc, data, base + k * stride))
    return out


def node_world(doc, ni, parent):
    n = doc["nodes"][ni]
    if "matrix" in n:
        M = np.array(n["matrix"], np.float64).reshape(4, 4).T
    else:
        M = np.eye(4)
        t = n.get("translation", [0, 0, 0])
        x, y, z, w = n.get("rotation", [0, 0, 0, 1])
        s = n.get("scale", [1, 1, 1])
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        M[:3, :3] = R * np.array(s)[None, :]
        M[:3, 3] = t
    return parent @ M


def scene_objects(path, shading=False):
    import base64
    doc = json.load(open(path))
    bufs = []
    for b in doc["buffers"]:
        uri = b["uri"]
        bufs.append(base64.b64decode(uri.split(",", 1)[1]) if uri.startswith("data:")
                    else open(os.path.join(os.path.dirname(path), uri), "rb").read())
    objs = []

    def visit(ni, parent):
        W = node_world(doc, ni, parent)
        n = doc["nodes"][ni]
        if "mesh" in n:
            for p in doc["meshes"][n["mesh"]]["primitives"]:
                assert p.get("mode", 4) == 4
                pos = np.array(read_accessor(doc, bufs, p["attributes"]["POSITION"]), np.float32)
                idx = np.array([t[0] for t in read_accessor(doc, bufs, p["indices"])], np.int64).reshape(-1, 3)
                if shading:
                    nrm = np.array(read_accessor(doc, bufs, p["attributes"]["NORMAL"]), np.float32)
                    mats = doc.get("materials", [])
                    m = mats[p["material"]] if "material" in p else {}
                    pbr = m.get("pbrMetallicRoughness", {})
                    mat = (np.array(pbr.get("baseColorFactor", [1, 1, 1, 1])[:3], np.float32),
                           np.array(m.get("emissiveFactor", [0, 0, 0]), np.float32),
                           np.array([pbr.get("metallicFactor", 1.0), pbr.get("roughnessFactor", 1.0)], np.float32))
                    objs.append((pos, idx, W, nrm, mat))
                else:
                    objs.append((pos, idx, W))
        for c in n.get("children", []):
            visit(c, W)

    for sc in doc.get("scenes", []):
        for r in sc["nodes"]:
            visit(r, np.eye(4))
    return objs


def main(media):
    rng = np.random.default_rng(1234)
    facts = {}
    for name, rel in [("cube", "cube.gltf"), ("mis_test", "mis_test/mis_test.gltf"), ("cbox", "cbox/cbox.gltf")]:
        objs = scene_objects(os.path.join(media, rel))
        facts[name] = [{"n_verts": int(p.shape[0]), "n_tris": int(f.shape[0]), "index_sum": int(f.sum()),
                        "index_checksum": int((f * np.arange(1, f.size + 1).reshape(f.shape)).sum()),
                        "pos_min": p.min(0).tolist(), "pos_max": p.max(0).tolist()} for p, f, _ in objs]
        Vw, Fw, base = [], [], 0
        for p, f, W in objs:
            Vw.append((p.astype(np.float64) @ W[:3, :3].T + W[:3, 3]).astype(np.float32))
            Fw.append(f + base)
            base += p.shape[0]
        Vw, Fw = np.concatenate(Vw), np.concatenate(Fw).astype(np.uint32)
        lo, hi = Vw.min(0).astype(np.float64), Vw.max(0).astype(np.float64)
        c, ext = (lo + hi) / 2, (hi - lo) * 0.6
        n = 2048
        o = rng.uniform(c - ext, c + ext, (n, 3))
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        R = np.zeros((n, 8), np.float32)
        R[:, :3], R[:, 3], R[:, 4:7], R[:, 7] = o, 1e-5, d, 1e7
        RA = R.copy()
        RA[:, 7] = rng.uniform(0.1, 2.0, n) * float(np.linalg.norm(hi - lo))
        Q = rng.uniform(c - ext, c + ext, (n, 3)).astype(np.float32)
        rp, rt, ruv = O.ray_closest_brute_f32(Vw, Fw, R)
        occ = O.ray_any_brute_f32(Vw, Fw, RA)
        qp, qpt, qd, qb = O.cpq_brute_f32(Vw, Fw, Q)
        xf = np.stack([W[:3].astype(np.float32).reshape(12) for _, _, W in objs])
        meshes = [(p, f.astype(np.uint32)) for p, f, _ in objs]
        io, ip, it, iuv = O.inst_ray_brute_f32(xf, np.arange(len(objs)), meshes, R)
        iocc = O.inst_ray_brute_f32(xf, np.arange(len(objs)), meshes, RA, any_hit=True)
        arrays = {"world_verts": Vw, "world_tris": Fw, "rays": R, "rays_any": RA, "queries": Q,
                  "ray_prim": rp, "ray_t": rt, "ray_uv": ruv, "occluded": occ,
                  "cpq_prim": qp, "cpq_pt": qpt, "cpq_dist": qd, "cpq_bary": qb,
                  "xf": xf, "inst_obj": io, "inst_prim": ip, "inst_t": it, "inst_uv": iuv, "inst_occluded": iocc}
        for k, (p, f, W) in enumerate(objs):
            arrays[f"obj{k}_pos"] = p
            arrays[f"obj{k}_tris"] = f.astype(np.uint32)
        np.savez_compressed(os.path.join(HERE, f"gltf_{name}.npz"), **arrays)
        print(name, len(objs), "objects", int(Fw.shape[0]), "triangles; hit rate", float((rp >= 0).mean()))
    with open(os.path.join(HERE, "gltf_facts.json"), "w") as f:
        json.dump(facts, f, indent=1)
    shading_fixtures(media)


# Cameras for rendering: cbox's own camera node (position (0, 0, 19.99), looking at the
# origin, yfov 0.7175 rad, aspect 4/3); mis_test has none, so a view of its four plates.
CAMERAS = {"cbox": {"eye": [0.0, 0.0, 19.992091970966317], "target": [0.0, 0.0, 0.0],
                    "yfov_deg": float(np.degrees(0.7175414562225342)), "aspect": 4.0 / 3.0},
           "mis_test": {"eye": [-2.5, 1.2, 0.0], "target": [0.5, 0.3, 0.0], "yfov_deg": 45.0, "aspect": 4.0 / 3.0}}


def shading_fixtures(media):
    """gltf_<name>_shading.npz: per object the vertex normals and the material factors
    (baseColorFactor rgb, emissiveFactor, metallic / roughness factors with the glTF
    defaults 1, 0, 1, 1), and a camera -- what the path tracer needs beyond gltf_<name>.npz."""
    for name, rel in [("mis_test", "mis_test/mis_test.gltf"), ("cbox", "cbox/cbox.gltf")]:
        objs = scene_objects(os.path.join(media, rel), shading=True)
        arrays = {"albedo": np.stack([o[4][0] for o in objs]), "emissive": np.stack([o[4][1] for o in objs]),
                  "metal_rough": np.stack([o[4][2] for o in objs])}
        for k, o in enumerate(objs):
            arrays[f"obj{k}_nrm"] = o[3]
        cam = CAMERAS[name]
        for key in ("eye", "target"):
            arrays["camera_" + key] = np.array(cam[key], np.float64)
        arrays["camera_yfov_deg"] = np.float64(cam["yfov_deg"])
        arrays["camera_aspect"] = np.float64(cam["aspect"])
        np.savez_compressed(os.path.join(HERE, f"gltf_{name}_shading.npz"), **arrays)
        print(name, "shading fixture:", len(objs), "objects,", int((arrays["emissive"] > 0).any(1).sum()), "lights")


if __name__ == "__main__":
    media = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/media"
    if len(sys.argv) > 2 and sys.argv[2] == "--shading-only":
        shading_fixtures(media)
    else:
        main(media)
