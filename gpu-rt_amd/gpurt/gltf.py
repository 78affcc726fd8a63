"""glTF 2.0 ingestion into the arrays the BVH builder takes.

Mirrors what the reference's loader extracts for its geometry (Scene::load and
Scene::parse_mesh, src/scene/scene.cpp:48-391):

* one object per mesh primitive, with the primitive's indices widened to uint32 from any
  integer component type (scene.cpp:60-107);
* TRIANGLE_FAN primitives rewritten as the triangle list (f0, f(i-1), f(i)) and
  TRIANGLE_STRIP primitives as (s(i-2), s(i-1), s(i)) for i >= 2, exactly the
  reference's rewrite (scene.cpp:109-141; note that it does not alternate the winding
  of strip triangles, and neither does this); points and lines are ignored
  (scene.cpp:278-283);
* POSITION, NORMAL (VEC3), TANGENT (VEC4) and TEXCOORD_0 (VEC2) from float or double
  accessors, packed per vertex as the 48-byte Mesh::Vertex
  {pos.xyz, u}, {normal.xyz, v}, {tangent} (scene.cpp:305-311; mesh.h:16-22), with
  absent attributes zero;
* the base colour, emissive and metallic/roughness factors of the material
  (scene.cpp:286-303).

Placement: each object carries its node's world matrix, the glTF product of parent and
child node matrices (column vectors; TRS nodes composed as T * R * S per the glTF
specification).  The reference's application-level placement policy (it multiplies by
the transposed node matrix, ignores TRS nodes, decomposes the result into a Pose and
applies a global scene scale, scene.cpp:353-370, pose.cpp:4-6) is not part of the
geometric path and is not reproduced.

    scene = gltf.load("cbox.gltf")
    for obj in scene["objects"]:
        obj["verts"]      # (n, 12) float32, Mesh::Vertex layout (48-byte stride)
        obj["indices"]    # (m, 3) uint32
        obj["transform"]  # (4, 4) float64 object-to-world
    tlas, blas = gltf.build_scene(ctx, scene)   # BLAS per object + TLAS over them
"""
import base64
import json
import os
import struct

import numpy as np

_COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5124: np.int32,
              5125: np.uint32, 5126: np.float32, 5130: np.float64}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}
_MODE_POINTS, _MODE_LINES, _MODE_LINE_LOOP, _MODE_LINE_STRIP = 0, 1, 2, 3
_MODE_TRIANGLES, _MODE_STRIP, _MODE_FAN = 4, 5, 6


def _read_glb(path):
    data = open(path, "rb").read()
    magic, version, length = struct.unpack_from("<III", data, 0)
    if magic != 0x46546C67 or version != 2:
        raise ValueError(f"{path}: not a glTF 2.0 binary")
    off, doc, bin_chunk = 12, None, None
    while off < length:
        clen, ctype = struct.unpack_from("<II", data, off)
        chunk = data[off + 8:off + 8 + clen]
        if ctype == 0x4E4F534A:  # JSON
            doc = json.loads(chunk.decode("utf-8"))
        elif ctype == 0x004E4942:  # BIN
            bin_chunk = chunk
        off += 8 + clen
    return doc, bin_chunk


def _buffers(doc, base_dir, bin_chunk):
    out = []
    for b in doc.get("buffers", []):
        uri = b.get("uri")
        if uri is None:
            if bin_chunk is None:
                raise ValueError("buffer without uri and no binary chunk")
            out.append(bin_chunk)
        elif uri.startswith("data:"):
            out.append(base64.b64decode(uri.split(",", 1)[1]))
        else:
            with open(os.path.join(base_dir, uri), "rb") as f:
                out.append(f.read())
    return out


def _accessor(doc, bufs, index):
    """(count, ncomp) array of an accessor, honouring byteStride; sparse accessors are
    not supported."""
    acc = doc["accessors"][index]
    if "sparse" in acc:
        raise ValueError("sparse accessors are not supported")
    dtype = np.dtype(_COMPONENT[acc["componentType"]])
    ncomp = _NCOMP[acc["type"]]
    count = acc["count"]
    if "bufferView" not in acc:
        return np.zeros((count, ncomp), dtype)
    view = doc["bufferViews"][acc["bufferView"]]
    data = bufs[view["buffer"]]
    start = view.get("byteOffset", 0) + acc.get("byteOffset", 0)
    elem = dtype.itemsize * ncomp
    stride = view.get("byteStride", 0) or elem
    if count == 0:
        return np.zeros((0, ncomp), dtype)
    raw = np.frombuffer(data, np.uint8, count=(count - 1) * stride + elem, offset=start)
    rows = np.lib.stride_tricks.as_strided(raw, shape=(count, elem), strides=(stride, 1))
    return np.ascontiguousarray(rows).view(dtype).reshape(count, ncomp)


def _triangle_list(idx, mode):
    """Index list -> (m, 3) triangles, as scene.cpp:109-141 rewrites fans and strips."""
    idx = idx.astype(np.int64).reshape(-1)
    if mode == _MODE_FAN:
        i = np.arange(2, idx.size)
        return np.stack([np.full(i.size, idx[0] if idx.size else 0), idx[i - 1], idx[i]], 1)
    if mode == _MODE_STRIP:
        i = np.arange(2, idx.size)
        return np.stack([idx[i - 2], idx[i - 1], idx[i]], 1)
    return idx[: idx.size - idx.size % 3].reshape(-1, 3)


def _node_matrix(node):
    if "matrix" in node:
        return np.asarray(node["matrix"], np.float64).reshape(4, 4).T  # column-major storage
    T = np.eye(4)
    T[:3, 3] = node.get("translation", [0.0, 0.0, 0.0])
    x, y, z, w = node.get("rotation", [0.0, 0.0, 0.0, 1.0])
    R = np.eye(4)
    R[:3, :3] = [[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                 [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                 [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]]
    S = np.diag(list(node.get("scale", [1.0, 1.0, 1.0])) + [1.0])
    return T @ R @ S


def _material(doc, index):
    mats = doc.get("materials", [])
    m = mats[index] if index is not None and 0 <= index < len(mats) else {}
    pbr = m.get("pbrMetallicRoughness", {})
    return {"albedo": np.asarray(pbr.get("baseColorFactor", [1, 1, 1, 1])[:3], np.float32),
            "emissive": np.asarray(m.get("emissiveFactor", [0, 0, 0]), np.float32),
            "metal_rough": np.asarray([pbr.get("metallicFactor", 1.0), pbr.get("roughnessFactor", 1.0)], np.float32),
            "name": m.get("name", "")}


def _primitive(doc, bufs, prim):
    mode = prim.get("mode", _MODE_TRIANGLES)
    if mode not in (_MODE_TRIANGLES, _MODE_STRIP, _MODE_FAN):
        return None  # points and lines are not triangle based
    attrs = prim["attributes"]
    if "POSITION" not in attrs:
        return None
    pos = _accessor(doc, bufs, attrs["POSITION"]).astype(np.float32)
    n = pos.shape[0]
    if "indices" in prim:
        idx = _accessor(doc, bufs, prim["indices"]).reshape(-1)
    else:
        idx = np.arange(n, dtype=np.int64)
    tris = _triangle_list(idx, mode)
    if tris.size and (tris.min() < 0 or tris.max() >= n):
        raise ValueError("triangle index outside the vertex range")
    verts = np.zeros((n, 12), np.float32)
    verts[:, 0:3] = pos
    if "NORMAL" in attrs:
        verts[:, 4:7] = _accessor(doc, bufs, attrs["NORMAL"])[:n, :3]
    if "TANGENT" in attrs:
        verts[:, 8:12] = _accessor(doc, bufs, attrs["TANGENT"])[:n, :4]
    if "TEXCOORD_0" in attrs:
        tc = _accessor(doc, bufs, attrs["TEXCOORD_0"])[:n, :2].astype(np.float32)
        verts[:, 3] = tc[:, 0]
        verts[:, 7] = tc[:, 1]
    return {"verts": verts, "indices": np.ascontiguousarray(tris, np.uint32),
            "material": _material(doc, prim.get("material"))}


def load(path):
    """Parse a .gltf (external or data: buffers) or .glb file.  Returns {"objects": [...]}
    with one object per triangle primitive of every mesh instanced by the scene graph,
    in scene / node / primitive order."""
    if path.lower().endswith(".glb"):
        doc, bin_chunk = _read_glb(path)
    else:
        with open(path) as f:
            doc = json.load(f)
        bin_chunk = None
    bufs = _buffers(doc, os.path.dirname(os.path.abspath(path)), bin_chunk)
    objects = []

    def visit(ni, parent):
        node = doc["nodes"][ni]
        world = parent @ _node_matrix(node)
        if "mesh" in node:
            mesh = doc["meshes"][node["mesh"]]
            for k, prim in enumerate(mesh.get("primitives", [])):
                obj = _primitive(doc, bufs, prim)
                if obj is None:
                    continue
                obj.update({"transform": world.copy(), "mesh": node["mesh"], "primitive": k,
                            "name": node.get("name", mesh.get("name", ""))})
                objects.append(obj)
        for c in node.get("children", []):
            visit(c, world)

    scenes = doc.get("scenes", [])
    roots = []
    if scenes:
        for sc in scenes:
            roots += sc.get("nodes", [])
    else:
        roots = list(range(len(doc.get("nodes", []))))
    for r in roots:
        visit(r, np.eye(4))
    return {"objects": objects}


def world_vertices(obj):
    """(n, 3) float32 world-space positions of an object (FP64 transform, then rounded)."""
    p = obj["verts"][:, :3].astype(np.float64)
    T = obj["transform"]
    return (p @ T[:3, :3].T + T[:3, 3]).astype(np.float32)


def build_scene(ctx, scene, **bvh_opts):
    """The reference's GPURT::build_accel (gpurt.cpp:220-241): one BLAS per object over
    its Mesh::Vertex array (48-byte stride, as Accel::recreate(Mesh) reads it) and a TLAS
    over the objects with their world matrices and custom index = object position.
    Returns (tlas, [bvh, ...])."""
    import torch
    import gpurt
    dev = ctx.device
    blas = []
    for obj in scene["objects"]:
        v = torch.from_numpy(obj["verts"]).to(dev)
        f = torch.from_numpy(obj["indices"].astype(np.int32)).to(dev)
        blas.append(gpurt.Bvh(ctx, v, f, **bvh_opts))
    tlas = gpurt.Tlas(ctx, [(b, o["transform"][:3], i) for i, (b, o) in enumerate(zip(blas, scene["objects"]))])
    return tlas, blas
