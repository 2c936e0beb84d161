"""glTF ingestion (vk_voxel_cone_tracing_amd.gltf) against known answers:
round trips of the procedural scenes through .gltf and .glb, a node
hierarchy with TRS transforms and a skipped camera subtree, the loader's
generated normals (unnormalized sums of face normals) and indices, its
cube-map texture coordinates, and the glTF material defaults."""
import base64
import json

import numpy as np
import pytest

from vk_voxel_cone_tracing_amd import gltf, scenes


@pytest.mark.parametrize("binary", [False, True])
def test_round_trip_of_procedural_scene(tmp_path, binary):
    sc = scenes.room(3, scale=2.0, n_boxes=8, n_spheres=1)
    path = tmp_path / ("room.glb" if binary else "room.gltf")
    gltf.save(sc, path, binary=binary)
    back = gltf.load(path)
    p0, n0, m0 = sc.world_triangles()
    p1, n1, m1 = back.world_triangles()
    np.testing.assert_array_equal(p0, p1)
    # the inverse-transpose is recomputed from the stored float world matrix
    np.testing.assert_allclose(n0, n1, rtol=1e-6, atol=1e-6)
    np.testing.assert_array_equal(back.materials["base_color_factor"][m1], sc.materials["base_color_factor"][m0])
    assert back.triangle_count == sc.triangle_count


def _doc_with(buffers_bytes: bytes, doc: dict) -> dict:
    doc["buffers"] = [{"byteLength": len(buffers_bytes),
                       "uri": "data:application/octet-stream;base64," + base64.b64encode(buffers_bytes).decode()}]
    doc["asset"] = {"version": "2.0"}
    return doc


def test_hierarchy_generated_normals_and_defaults(tmp_path):
    # mesh 0: a unit quad in the xy plane, u16 indices, no normals / uvs
    quad = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    qidx = np.array([0, 1, 2, 0, 2, 3], np.uint16)
    # mesh 1: one triangle, no indices, positions interleaved with a pad (byteStride 16)
    tri = np.array([[0, 0, 0, 9], [0, 0, 1, 9], [0, 1, 0, 9]], np.float32)
    blob = quad.tobytes() + qidx.tobytes() + b"\0\0\0\0" + tri.tobytes()
    views = [{"buffer": 0, "byteOffset": 0, "byteLength": 48},
             {"buffer": 0, "byteOffset": 48, "byteLength": 12},
             {"buffer": 0, "byteOffset": 64, "byteLength": 48, "byteStride": 16}]
    acc = [{"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3", "min": [0, 0, 0], "max": [1, 1, 0]},
           {"bufferView": 1, "componentType": 5123, "count": 6, "type": "SCALAR"},
           {"bufferView": 2, "componentType": 5126, "count": 3, "type": "VEC3", "min": [0, 0, 0], "max": [0, 1, 1]}]
    s = np.sqrt(0.5)
    doc = _doc_with(blob, {
        "scene": 0, "scenes": [{"nodes": [0, 3]}],
        "nodes": [
            # parent: translate (10,0,0), rotate 90 deg about +y, scale 2; no mesh
            {"translation": [10, 0, 0], "rotation": [0, s, 0, s], "scale": [2, 2, 2], "children": [1, 2]},
            {"mesh": 0, "translation": [0, 0, 5]},
            {"mesh": 1},
            {"camera": 0, "children": [4]},  # camera subtree: skipped
            {"mesh": 0},
        ],
        "cameras": [{"type": "perspective", "perspective": {"yfov": 1.0, "znear": 0.1}}],
        "meshes": [{"primitives": [{"attributes": {"POSITION": 0}, "indices": 1, "material": 1}]},
                   {"primitives": [{"attributes": {"POSITION": 2}}]}],
        "materials": [{"name": "plain"},
                      {"pbrMetallicRoughness": {"baseColorFactor": [0.5, 0.25, 1.0, 1.0], "metallicFactor": 0.0,
                                                "roughnessFactor": 0.3}, "emissiveFactor": [1, 0, 0]}],
        "accessors": acc, "bufferViews": views})
    path = tmp_path / "h.gltf"
    path.write_text(json.dumps(doc))
    sc = gltf.load(path)
    assert sc.primitives.shape == (2, 5)  # two mesh nodes; the camera's child is not visited
    # world = T(10,0,0) Ry(90) S(2) [T(0,0,5)] ; Ry(90): x -> -z, z -> x
    def world(p, child_t):
        q = np.asarray(p, np.float64) + child_t
        q = 2.0 * q
        q = np.array([q[2], q[1], -q[0]])
        return q + [10, 0, 0]

    pos, nrm, mat = sc.world_triangles()
    want0 = np.array([[world(quad[i], [0, 0, 5]) for i in t] for t in (qidx[:3], qidx[3:])]).reshape(2, 9)
    np.testing.assert_allclose(pos[:2], want0, atol=1e-5)
    want1 = np.array([world(tri[i, :3], [0, 0, 0]) for i in range(3)]).reshape(9)
    np.testing.assert_allclose(pos[2], want1, atol=1e-5)
    # generated normals: vertices 0 and 2 of the quad are shared by both triangles -> (0,0,2)
    np.testing.assert_allclose(sc.normals[:4], [[0, 0, 2], [0, 0, 1], [0, 0, 2], [0, 0, 1]], atol=1e-6)
    # the second mesh has no indices: 0..2; its triangle (0,0,0),(0,0,1),(0,1,0) faces -x
    np.testing.assert_array_equal(sc.indices[6:], [0, 1, 2])
    np.testing.assert_allclose(sc.normals[4:7], np.tile([-1.0, 0.0, 0.0], (3, 1)), atol=1e-6)
    # materials: defaults for "plain", the given factors for the other
    m0, m1 = sc.materials[0], sc.materials[1]
    assert list(m0["base_color_factor"]) == [1, 1, 1, 1] and m0["metallic_factor"] == 1.0
    assert m0["roughness_factor"] == 1.0 and m0["alpha_cutoff"] == 0.5 and m0["base_color_texture"] == -1
    assert m1["roughness_factor"] == np.float32(0.3) and list(m1["emissive_factor"]) == [1, 0, 0]
    assert mat[0] == 1 and mat[2] == 0
    # scene box: accessor min/max through the world matrices
    corners = [world(c, [0, 0, 5]) for c in ([0, 0, 0], [1, 1, 0])] + [world(c, [0, 0, 0]) for c in ([0, 0, 0], [0, 1, 1])]
    np.testing.assert_allclose(sc.bb_min, np.min(corners, 0), atol=1e-5)
    np.testing.assert_allclose(sc.bb_max, np.max(corners, 0), atol=1e-5)


def test_cubemap_texcoords():
    uv = gltf._cubemap_uv(np.array([[1.0, 0.5, 0.2], [-2.0, 1.0, 0.5], [0.2, 1.0, -0.4], [0.3, -0.2, -1.0]],
                                   np.float32))
    np.testing.assert_allclose(uv[0], [0.4, 0.75])       # +x: u from -z, v from y
    np.testing.assert_allclose(uv[1], [0.625, 0.75])     # -x: u from z
    np.testing.assert_allclose(uv[2], [0.6, 0.7])        # +y: u from x, v from -z
    np.testing.assert_allclose(uv[3], [0.35, 0.4])       # -z: u from -x
    # the loader's +z case tests |y| instead of |z|: a +z-dominant vertex gets
    # no projection axis, and its coordinates divide by zero
    assert not np.all(np.isfinite(gltf._cubemap_uv(np.array([[0.1, 0.2, 1.0]], np.float32))))
