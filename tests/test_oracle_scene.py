"""Pins of the oracle's input producers (oracle/orc_scene.c) by analytic
scenes: a plane facing a perspective camera has NDC depth
far / (far - near) - far * near / ((far - near) * z) at view distance z
(glm's perspectiveLH_ZO), and an orthographic light sees a horizontal quad at
NDC depth (distance - near) / (far - near) (orthoLH_ZO), with texels whose
centres miss the quad reading 1.0. G-buffer channels: base colour and
roughness bytes, normal * 0.5 + 0.5 in halves."""
import ctypes as C

import numpy as np

import oracle_api as O
from vk_voxel_cone_tracing_amd import glm


def _quad(corners, normal):
    a, b, c, d = [np.asarray(p, np.float32) for p in corners]
    pos = np.concatenate([a, b, c, a, c, d]).astype(np.float32)
    nrm = np.tile(np.asarray(normal, np.float32), 6)
    return pos, nrm


def _mats(base, emissive=(0.0, 0.0, 0.0)):
    arr = (O.Material * 1)()
    arr[0].base_color = (C.c_float * 4)(*base)
    arr[0].emissive = (C.c_float * 3)(*emissive)
    return arr


def test_gbuffer_of_a_plane_facing_the_camera():
    W, H = 16, 12
    near, far, z = 0.1, 100.0, 5.0
    pos, nrm = _quad([(-10, -10, z), (10, -10, z), (10, 10, z), (-10, 10, z)], (0, 0, -1))
    mat = np.zeros(2, np.uint32)
    view = glm.look_at_lh([0, 0, 0], [0, 0, 1], [0, 1, 0])
    proj = glm.perspective_lh_zo(np.radians(60.0), W / H, near, far)
    vp = proj @ view
    g = dict(diffuse=np.zeros((H, W), np.uint32), normal=np.zeros((H, W, 4), np.uint16),
             specular=np.zeros((H, W), np.uint32), emission=np.zeros((H, W, 4), np.uint16),
             depth=np.zeros((H, W), np.float32))
    surf = np.array([0.4, 0.7, 0.2, 0.6, 1.0], np.float32)  # roughness, metallic, specular rgb
    t = O.tris_struct(pos, nrm, mat)
    O.lib().orc_raycast_gbuffer(C.byref(t), _mats((0.25, 0.5, 1.0, 1.0)), O.fptr(surf),
                                O.fptr(glm.flat(np.linalg.inv(vp))), O.fptr(glm.flat(vp)), W, H,
                                O.vptr(g["diffuse"]), O.vptr(g["normal"]), O.vptr(g["specular"]),
                                O.vptr(g["emission"]), O.vptr(g["depth"]))
    want = far / (far - near) - far * near / ((far - near) * z)
    assert np.allclose(g["depth"], want, rtol=0, atol=2e-6)
    # bytes: round(0.25 * 255) = 64, 128, 255, roughness round(0.4 * 255) = 102
    assert np.all(g["diffuse"] == (64 | 128 << 8 | 255 << 16 | 102 << 24))
    # specular rgb round(0.2, 0.6, 1.0 * 255) = 51, 153, 255; metallic 0.7f = 0.69999999 -> 178.49999 -> 178
    assert np.all(g["specular"] == (51 | 153 << 8 | 255 << 16 | 178 << 24))
    half = [O.lib().orc_float_to_half(v) for v in (0.5, 0.5, 0.0, 1.0)]
    assert np.all(g["normal"] == np.array(half, np.uint16))


def test_gbuffer_background_is_far():
    W, H = 8, 6
    pos, nrm = _quad([(-1, -1, -5), (1, -1, -5), (1, 1, -5), (-1, 1, -5)], (0, 0, 1))  # behind the camera
    view = glm.look_at_lh([0, 0, 0], [0, 0, 1], [0, 1, 0])
    vp = glm.perspective_lh_zo(np.radians(60.0), W / H, 0.1, 100.0) @ view
    g = O.raycast_gbuffer(type("S", (), dict(materials=np.zeros(1, dtype=[("base_color_factor", "f4", 4),
                                                                        ("emissive_factor", "f4", 3)]),
                                             surface=np.zeros((1, 5), np.float32)))(),
                          pos, nrm, np.zeros(2, np.uint32), vp, np.linalg.inv(vp), W, H)
    assert np.all(g["depth"] == 1.0)
    assert np.all(g["diffuse"] == 0)


def test_shadow_map_of_a_floor_and_a_raised_quad():
    W = H = 64
    near, far = 0.1, 30.0
    floor = _quad([(-20, 0, -20), (20, 0, -20), (20, 0, 20), (-20, 0, 20)], (0, 1, 0))
    top = _quad([(-3.3, 2, -3.3), (2.7, 2, -3.3), (2.7, 2, 2.7), (-3.3, 2, 2.7)], (0, 1, 0))
    pos = np.concatenate([floor[0], top[0]])
    nrm = np.concatenate([floor[1], top[1]])
    mat = np.zeros(4, np.uint32)
    eye = np.array([0.0, 10.0, 0.0])
    light = dict(view=glm.look_at_lh(eye, eye + [0.0, -1.0, 0.0], [0.0, 0.0, 1.0]),
                 proj=glm.ortho_lh_zo(-16, 16, -16, 16, near, far))
    depth = O.raycast_shadow(pos, nrm, mat, light, W, H)
    # view x = world x, view y = world z; texel centres at -16 + 0.5 * i + 0.25
    idx = np.arange(W)
    centre = -16.0 + 0.5 * idx + 0.25
    inside = (centre >= -3.3) & (centre <= 2.7)
    assert inside.sum() == 12 and idx[inside][0] == 25
    want = np.where(inside[:, None] & inside[None, :], (8.0 - near) / (far - near), (10.0 - near) / (far - near))
    assert np.allclose(depth, want, rtol=0, atol=2e-6)
