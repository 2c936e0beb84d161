"""glTF 2.0 ingestion into the reference's mesh streams.

Follows GLTFLoader::loadScene / processMesh / processNode / importMaterials
(VFS/Util/GLTFLoader.cpp:43-126, 128-393, 407-486, 560-587) and
GLTFScene::uploadMatrixBuffer / cmdDraw (VFS/GLTFScene.cpp:398-409, 457-489):

* every triangle primitive of every mesh becomes one primitive mesh with its
  own vertex range (vertex_offset) and index range (first_index, indices local
  to the primitive); primitives without indices get 0..n-1;
* missing NORMAL: per vertex the sum of the unit normals of its triangles
  (normalize(cross(normalize(e0), normalize(e1))), the sum left unnormalized,
  GLTFLoader.cpp:189-211);
* missing TEXCOORD_0: the loader's cube-map projection (GLTFLoader.cpp:215-290),
  its positive-z test included as written (it compares |y|, not |z|);
* the default scene's node tree is flattened depth first; a node's local
  matrix is T * R * S * matrix and its world matrix parent.world * local
  (GetLocalMatrix, :510-516); camera and light nodes are not mesh nodes and
  their children are not visited (:441-480);
* one (world, transpose(inverse(world))) pair per mesh node; one draw per
  (mesh node, primitive) with instance = the mesh node's ordinal. (The
  reference's cmdDraw advances its instance index for mesh-less nodes too,
  while its matrix buffer holds mesh nodes only; with mesh-less nodes in the
  tree the two disagree. Here the instance index is the mesh node's ordinal.)
* materials with the glTF defaults tinygltf fills in;
* the scene box from each primitive's accessor min/max transformed by its
  node's world matrix, as calculateSceneDimension does (:517-540);
* textures: texture i is the image of glTF texture i's source (:118-123),
  decoded to RGBA8 as tinygltf's stb_image loader delivers it (RGB gets
  alpha 255, grey is replicated, 16-bit channels keep their high byte).
  Images are decoded with Pillow (PNG decoding is lossless, so identical;
  a JPEG decoder may differ from stb_image's by a count in some texels).

Supports .gltf (external .bin files or base64 data URIs) and .glb.
"""
from __future__ import annotations

import base64
import json
import struct
from pathlib import Path

import numpy as np

from .scenes import MATERIAL_DTYPE, Scene

_COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}


def _load_json_and_buffers(path: Path):
    data = path.read_bytes()
    if data[:4] == b"glTF":
        _, _, total = struct.unpack_from("<III", data, 0)
        off = 12
        doc, bin_chunk = None, None
        while off < total:
            ln, kind = struct.unpack_from("<II", data, off)
            chunk = data[off + 8: off + 8 + ln]
            if kind == 0x4E4F534A:
                doc = json.loads(chunk.decode("utf-8"))
            elif kind == 0x004E4942:
                bin_chunk = chunk
            off += 8 + ln
        buffers = []
        for i, b in enumerate(doc.get("buffers", [])):
            buffers.append(bin_chunk if ("uri" not in b and i == 0) else _read_uri(path, b["uri"]))
        return doc, buffers
    doc = json.loads(data.decode("utf-8"))
    return doc, [_read_uri(path, b["uri"]) for b in doc.get("buffers", [])]


def _read_uri(path: Path, uri: str) -> bytes:
    if uri.startswith("data:"):
        return base64.b64decode(uri.split(",", 1)[1])
    return (path.parent / uri).read_bytes()


def _accessor(doc, buffers, idx) -> np.ndarray:
    acc = doc["accessors"][idx]
    dt = np.dtype(_COMPONENT[acc["componentType"]])
    nc = _NCOMP[acc["type"]]
    count = acc["count"]
    if "bufferView" not in acc:
        return np.zeros((count, nc) if nc > 1 else count, dt)
    bv = doc["bufferViews"][acc["bufferView"]]
    buf = buffers[bv["buffer"]]
    start = bv.get("byteOffset", 0) + acc.get("byteOffset", 0)
    stride = bv.get("byteStride", 0) or dt.itemsize * nc
    if stride == dt.itemsize * nc:
        a = np.frombuffer(buf, dt, count * nc, start).reshape(count, nc)
    else:
        rows = [np.frombuffer(buf, dt, nc, start + r * stride) for r in range(count)]
        a = np.stack(rows) if rows else np.zeros((0, nc), dt)
    return a if nc > 1 else a.reshape(count)


def _decode_image(data: bytes) -> np.ndarray:
    """RGBA8 texels (H, W) uint32, R in the low byte."""
    import io

    try:
        from PIL import Image
    except ImportError as e:  # pragma: no cover - Pillow ships with this image
        raise RuntimeError("decoding glTF images needs Pillow") from e
    im = Image.open(io.BytesIO(data))
    if im.mode in ("I;16", "I;16B", "I;16L", "I"):
        g = (np.asarray(im, np.uint32) >> 8).astype(np.uint8)  # stb_image keeps the high byte
        rgba = np.stack([g, g, g, np.full_like(g, 255)], -1)
    else:
        rgba = np.asarray(im.convert("RGBA"), np.uint8)
    r = rgba.astype(np.uint32)
    return np.ascontiguousarray(r[..., 0] | (r[..., 1] << 8) | (r[..., 2] << 16) | (r[..., 3] << 24))


def _image_bytes(path: Path, doc, buffers, img: dict) -> bytes:
    if "uri" in img:
        return _read_uri(path, img["uri"])
    bv = doc["bufferViews"][img["bufferView"]]
    start = bv.get("byteOffset", 0)
    return bytes(buffers[bv["buffer"]][start: start + bv["byteLength"]])


def _encode_png(texels: np.ndarray) -> bytes:
    import io

    from PIL import Image

    t = np.ascontiguousarray(texels, np.uint32)
    rgba = np.stack([(t >> s) & 0xFF for s in (0, 8, 16, 24)], -1).astype(np.uint8)
    out = io.BytesIO()
    Image.fromarray(rgba, "RGBA").save(out, format="PNG")
    return out.getvalue()


def _quat_to_mat4(q):
    # glm::toMat4 of quat(w, x, y, z) with q = [x, y, z, w]
    x, y, z, w = [float(v) for v in q]
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y), 0.0],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x), 0.0],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y), 0.0],
        [0.0, 0.0, 0.0, 1.0],
    ])


def local_matrix(node: dict) -> np.ndarray:
    """GetLocalMatrix: translate(T) * toMat4(R) * scale(S) * matrix (glm
    column-major storage of node.matrix)."""
    T = np.eye(4)
    T[:3, 3] = node.get("translation", [0.0, 0.0, 0.0])
    R = _quat_to_mat4(node.get("rotation", [0.0, 0.0, 0.0, 1.0]))
    S = np.diag(list(node.get("scale", [1.0, 1.0, 1.0])) + [1.0])
    M = np.array(node["matrix"], np.float64).reshape(4, 4).T if "matrix" in node else np.eye(4)
    return T @ R @ S @ M


def _cubemap_uv(p: np.ndarray) -> np.ndarray:
    out = np.empty((len(p), 2), np.float32)
    for k, (x, y, z) in enumerate(p.astype(np.float32)):
        ax, ay, az = abs(x), abs(y), abs(z)
        m, uc, vc = np.float32(0.0), np.float32(0.0), np.float32(0.0)
        if x > 0 and ax >= ay and ax >= az:
            m, uc, vc = ax, -z, y
        if not x > 0 and ax >= ay and ax >= az:
            m, uc, vc = ax, z, y
        if y > 0 and ay >= ax and ay >= az:
            m, uc, vc = ay, x, -z
        if not y > 0 and ay >= ax and ay >= az:
            m, uc, vc = ay, x, z
        if z > 0 and ay >= ax and ay >= az:  # the loader's positive-z test, as written
            m, uc, vc = az, x, y
        if not z > 0 and az >= ax and az >= ay:
            m, uc, vc = az, -x, y
        with np.errstate(divide="ignore", invalid="ignore"):
            out[k] = [(np.float32(uc) / np.float32(m) + np.float32(1.0)) * np.float32(0.5),
                      (np.float32(vc) / np.float32(m) + np.float32(1.0)) * np.float32(0.5)]
    return out


def _vertex_normals(pos: np.ndarray, idx: np.ndarray) -> np.ndarray:
    n = np.zeros_like(pos, dtype=np.float32)
    for t in range(0, len(idx) - 2, 3):
        i0, i1, i2 = int(idx[t]), int(idx[t + 1]), int(idx[t + 2])
        e0 = pos[i1] - pos[i0]
        e1 = pos[i2] - pos[i0]
        e0 = e0 / np.linalg.norm(e0)
        e1 = e1 / np.linalg.norm(e1)
        c = np.cross(e0, e1)
        c = c / np.linalg.norm(c)
        n[i0] += c
        n[i1] += c
        n[i2] += c
    return n


def load(path) -> Scene:
    """A Scene (the reference's mesh streams) from a .gltf or .glb file."""
    path = Path(path)
    doc, buffers = _load_json_and_buffers(path)
    positions, normals, texcoords, indices = [], [], [], []
    prim_meshes = []  # (first_index, index_count, vertex_offset, material, min, max)
    mesh_prims = []
    nvert = nidx = 0
    for mesh in doc.get("meshes", []):
        ids = []
        for prim in mesh.get("primitives", []):
            if prim.get("mode", 4) != 4:
                continue
            attrs = prim["attributes"]
            pos = _accessor(doc, buffers, attrs["POSITION"]).astype(np.float32)
            if "indices" in prim:
                idx = _accessor(doc, buffers, prim["indices"]).astype(np.uint32)
            else:
                idx = np.arange(len(pos), dtype=np.uint32)
            nrm = (_accessor(doc, buffers, attrs["NORMAL"]).astype(np.float32) if "NORMAL" in attrs
                   else _vertex_normals(pos, idx))
            uv = (_accessor(doc, buffers, attrs["TEXCOORD_0"]).astype(np.float32) if "TEXCOORD_0" in attrs
                  else _cubemap_uv(pos))
            acc = doc["accessors"][attrs["POSITION"]]
            mn = np.array(acc.get("min", [0.0, 0.0, 0.0]), np.float32)
            mx = np.array(acc.get("max", [0.0, 0.0, 0.0]), np.float32)
            ids.append(len(prim_meshes))
            prim_meshes.append((nidx, len(idx), nvert, max(prim.get("material", 0), 0), mn, mx))
            positions.append(pos)
            normals.append(nrm)
            texcoords.append(uv)
            indices.append(idx)
            nvert += len(pos)
            nidx += len(idx)
        mesh_prims.append(ids)
    # flatten the default scene (processNode)
    worlds, draws = [], []

    def visit(ni, parent_world):
        node = doc["nodes"][ni]
        world = parent_world @ local_matrix(node)
        if "camera" in node or "KHR_lights_punctual" in node.get("extensions", {}):
            return
        if "mesh" in node and mesh_prims[node["mesh"]]:
            inst = len(worlds)
            worlds.append(world)
            for pm in mesh_prims[node["mesh"]]:
                draws.append((pm, inst))
        for c in node.get("children", []):
            visit(c, world)

    scenes = doc.get("scenes", [{"nodes": list(range(len(doc.get("nodes", []))))}])
    default = doc.get("scene", 0)
    for ni in scenes[default if default >= 0 else 0].get("nodes", []):
        visit(ni, np.eye(4))
    nodes = np.zeros((len(worlds), 32), np.float32)
    for i, w in enumerate(worlds):
        it = np.linalg.inv(w).T  # transpose(inverse(world))
        nodes[i, :16] = w.T.reshape(16)  # column-major storage
        nodes[i, 16:] = it.T.reshape(16)
    prims = np.array([[prim_meshes[pm][0], prim_meshes[pm][1], prim_meshes[pm][2], inst, prim_meshes[pm][3]]
                      for pm, inst in draws], np.uint32).reshape(-1, 5)
    mats = doc.get("materials", [])
    materials = np.zeros(max(1, len(mats)), MATERIAL_DTYPE)
    surface = np.zeros((max(1, len(mats)), 5), np.float32)
    for i in range(len(materials)):
        m = mats[i] if i < len(mats) else {}
        pbr = m.get("pbrMetallicRoughness", {})
        materials[i]["base_color_factor"] = pbr.get("baseColorFactor", [1.0, 1.0, 1.0, 1.0])
        materials[i]["base_color_texture"] = pbr.get("baseColorTexture", {}).get("index", -1)
        materials[i]["metallic_factor"] = pbr.get("metallicFactor", 1.0)
        materials[i]["roughness_factor"] = pbr.get("roughnessFactor", 1.0)
        materials[i]["metallic_roughness_texture"] = pbr.get("metallicRoughnessTexture", {}).get("index", -1)
        materials[i]["emissive_texture"] = m.get("emissiveTexture", {}).get("index", -1)
        mode = m.get("alphaMode", "OPAQUE")
        materials[i]["alpha_mode"] = 1 if mode == "MASK" else (2 if mode == "BLEND" else 0)
        materials[i]["alpha_cutoff"] = m.get("alphaCutoff", 0.5)
        materials[i]["double_sided"] = 1 if m.get("doubleSided", False) else 0
        materials[i]["emissive_factor"] = m.get("emissiveFactor", [0.0, 0.0, 0.0])
        materials[i]["normal_texture"] = m.get("normalTexture", {}).get("index", -1)
        materials[i]["normal_texture_scale"] = m.get("normalTexture", {}).get("scale", 1.0)
        materials[i]["occlusion_texture"] = m.get("occlusionTexture", {}).get("index", -1)
        materials[i]["occlusion_texture_strength"] = m.get("occlusionTexture", {}).get("strength", 1.0)
        base = np.asarray(materials[i]["base_color_factor"][:3], np.float32)
        metal = float(materials[i]["metallic_factor"])
        spec = 0.04 * (1.0 - metal) + base * metal
        surface[i] = [materials[i]["roughness_factor"], metal, *spec]
    # calculateSceneDimension: min and max corners through each node's world matrix
    lo = np.full(3, np.finfo(np.float32).max)
    hi = np.full(3, np.finfo(np.float32).tiny)
    for pm, inst in draws:
        w = worlds[inst]
        for corner in (prim_meshes[pm][4], prim_meshes[pm][5]):
            p = (w @ np.append(corner.astype(np.float64), 1.0))[:3]
            lo = np.minimum(lo, p)
            hi = np.maximum(hi, p)
    if np.array_equal(lo, hi) or not draws:
        lo, hi = np.full(3, -1.0), np.full(3, 1.0)
    textures = [_decode_image(_image_bytes(path, doc, buffers, doc["images"][t["source"]]))
                for t in doc.get("textures", [])]
    cat = lambda xs, shape, dt: np.concatenate(xs).astype(dt) if xs else np.zeros(shape, dt)  # noqa: E731
    return Scene(positions=cat(positions, (0, 3), np.float32), normals=cat(normals, (0, 3), np.float32),
                 texcoords=cat(texcoords, (0, 2), np.float32), indices=cat(indices, (0,), np.uint32), nodes=nodes,
                 materials=materials, primitives=prims, surface=surface, bb_min=lo.astype(np.float32),
                 bb_max=hi.astype(np.float32), textures=textures)


def save(scene: Scene, path, binary: bool = False) -> None:
    """Writes a Scene as glTF 2.0 (one mesh per primitive, one node per
    instance with its world matrix). Used to make test assets."""
    path = Path(path)
    blob = bytearray()
    views, accessors, meshes = [], [], []

    def add(arr, target=None):
        arr = np.ascontiguousarray(arr)
        while len(blob) % 4:
            blob.append(0)
        off = len(blob)
        blob.extend(arr.tobytes())
        v = {"buffer": 0, "byteOffset": off, "byteLength": arr.nbytes}
        if target:
            v["target"] = target
        views.append(v)
        comp = {np.dtype(np.float32): 5126, np.dtype(np.uint32): 5125, np.dtype(np.uint16): 5123}[arr.dtype]
        typ = {1: "SCALAR", 2: "VEC2", 3: "VEC3"}[1 if arr.ndim == 1 else arr.shape[1]]
        a = {"bufferView": len(views) - 1, "componentType": comp, "count": len(arr), "type": typ}
        if typ == "VEC3" and comp == 5126:
            a["min"] = arr.min(0).tolist()
            a["max"] = arr.max(0).tolist()
        accessors.append(a)
        return len(accessors) - 1

    nodes = []
    for first, count, voff, inst, mat in scene.primitives:
        idx = scene.indices[first: first + count]
        vmax = int(idx.max()) + 1 if len(idx) else 0
        sl = slice(int(voff), int(voff) + vmax)
        prim = {"attributes": {"POSITION": add(scene.positions[sl], 34962), "NORMAL": add(scene.normals[sl], 34962),
                               "TEXCOORD_0": add(scene.texcoords[sl], 34962)},
                "indices": add(idx.astype(np.uint32), 34963), "material": int(mat), "mode": 4}
        meshes.append({"primitives": [prim]})
        nodes.append({"mesh": len(meshes) - 1, "matrix": [float(x) for x in scene.nodes[inst, :16]]})
    mats = []
    for m in scene.materials:
        mat = {"pbrMetallicRoughness": {"baseColorFactor": [float(x) for x in m["base_color_factor"]],
                                        "metallicFactor": float(m["metallic_factor"]),
                                        "roughnessFactor": float(m["roughness_factor"])},
               "emissiveFactor": [float(x) for x in m["emissive_factor"]]}
        if m["base_color_texture"] >= 0:
            mat["pbrMetallicRoughness"]["baseColorTexture"] = {"index": int(m["base_color_texture"])}
        if m["emissive_texture"] >= 0:
            mat["emissiveTexture"] = {"index": int(m["emissive_texture"])}
        if m["occlusion_texture"] >= 0:
            mat["occlusionTexture"] = {"index": int(m["occlusion_texture"])}
        mats.append(mat)
    images = []
    for t in getattr(scene, "textures", []):
        png = _encode_png(t)
        while len(blob) % 4:
            blob.append(0)
        views.append({"buffer": 0, "byteOffset": len(blob), "byteLength": len(png)})
        blob.extend(png)
        images.append({"bufferView": len(views) - 1, "mimeType": "image/png"})
    doc = {"asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": list(range(len(nodes)))}], "nodes": nodes,
           "meshes": meshes, "materials": mats, "accessors": accessors, "bufferViews": views,
           "buffers": [{"byteLength": len(blob)}]}
    if images:
        # REPEAT / LINEAR, the sampler the reference creates for every texture
        doc["samplers"] = [{"magFilter": 9729, "minFilter": 9729, "wrapS": 10497, "wrapT": 10497}]
        doc["images"] = images
        doc["textures"] = [{"source": i, "sampler": 0} for i in range(len(images))]
    if binary:
        js = json.dumps(doc).encode("utf-8")
        js += b" " * (-len(js) % 4)
        bb = bytes(blob) + b"\0" * (-len(blob) % 4)
        total = 12 + 8 + len(js) + 8 + len(bb)
        path.write_bytes(struct.pack("<III", 0x46546C67, 2, total) + struct.pack("<II", len(js), 0x4E4F534A) + js +
                         struct.pack("<II", len(bb), 0x004E4942) + bb)
    else:
        doc["buffers"][0]["uri"] = path.with_suffix(".bin").name
        path.with_suffix(".bin").write_bytes(bytes(blob))
        path.write_text(json.dumps(doc))
