"""Pins of the octree oracle (oracle/orc_svo.c, orc_trace.c) that do not need
the reference to run:

- fragment words (voxelizer.frag:99-100) against hand-packed known answers,
  and the decode of octreeNodeFlag.comp:28-32 as the exact inverse on every
  12-bit coordinate of each axis;
- the leaf running average and the D5 mip-level selection
  (octreeNodeLeafWrite.comp:73-89, octreeNodeMipmapWrite.comp:41-52) on a
  hand-built 16^3 tree whose node colours were worked out by hand;
- sampleSVO (voxelConeTracing_Octree.frag:319-346) on a hand-built pool:
  slot order z | x << 1 | y << 2, the box-to-cell mapping, the early stop at
  an empty node;
- the octree trace (K16) on an empty pool: ambient occlusion
  clamp(0.5 * (1 + S) / S) with S the sum of the cosines of the cones in the
  normal's hemisphere, the octree pass dividing by that sum (D7).
"""
import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

import oracle_api as O

ROOT = Path(__file__).resolve().parents[1]
HALF = 1 << 3        # ORC_QUIRK_SVO_HALF_COORDS
MIP_TEST = 1 << 4    # ORC_QUIRK_SVO_MIP_LEVEL_TEST


def _lib():
    L = O.lib()
    L.orc_svo_pack_fragments.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
    L.orc_svo_unpack_fragments.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
    L.orc_sample_svo.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                 C.c_void_p]
    return L


def _pack(xyz, rgba):
    xyz = np.ascontiguousarray(xyz, np.uint32).reshape(-1, 3)
    rgba = np.ascontiguousarray(rgba, np.uint32).reshape(-1)
    out = np.zeros((len(rgba), 2), np.uint32)
    _lib().orc_svo_pack_fragments(xyz.ctypes.data, rgba.ctypes.data, len(rgba), out.ctypes.data)
    return out


def _unpack(frags, quirks=0):
    frags = np.ascontiguousarray(frags, np.uint32).reshape(-1, 2)
    out = np.zeros((len(frags), 3), np.uint32)
    _lib().orc_svo_unpack_fragments(frags.ctypes.data, len(frags), quirks, out.ctypes.data)
    return out


def test_fragment_words_known_answers():
    # x = 1, y = 2, z = 0x3ff: word0 = 1 | 2 << 12 | 0xff << 24, word1 = 3 << 28 | colour & 0x0fffffff
    f = _pack([[1, 2, 0x3FF], [4095, 0, 4095], [0, 4095, 256]], [0xFF8040C0, 0x12345678, 0xF0000000])
    assert f[0].tolist() == [0xFF002001, 0x3F8040C0]
    assert f[1].tolist() == [0xFF000FFF, 0xF2345678]
    assert f[2].tolist() == [0x00FFF000, 0x10000000]


def test_fragment_decode_inverts_every_12_bit_coordinate():
    v = np.arange(4096, dtype=np.uint32)
    others = [0, 1, 255, 256, 2047, 4095]
    for axis in range(3):
        for o in others:
            xyz = np.full((4096, 3), o, np.uint32)
            xyz[:, axis] = v
            f = _pack(xyz, np.full(4096, 0xFFFFFFFF, np.uint32))
            assert np.array_equal(_unpack(f), xyz)
            assert np.array_equal(_unpack(f, HALF), xyz >> 1)
            # the colour keeps 28 bits: alpha's high nibble is dropped (D6)
            assert np.all((f[:, 1] & 0x0FFFFFFF) == 0x0FFFFFFF)


def _node(pool, cell, depth, level):
    """Pool index of the node covering `cell` (leaf coordinates at 2^level)
    at `depth` (1 = root child), following child pointers (slot
    z | x << 1 | y << 2)."""
    idx, cur = 0, 0
    res = 1 << level
    p = list(cell)
    for _ in range(depth):
        res >>= 1
        c = [int(p[k] >= res) for k in range(3)]
        idx = cur + (c[2] | (c[0] << 1) | (c[1] << 2))
        cur = int(pool[idx, 0]) & 0x7FFFFFFF
        p = [p[k] - c[k] * res for k in range(3)]
    return idx


def _rgba(word):
    return [(int(word) >> (8 * k)) & 0xFF for k in range(4)]


def _build(frags, level, quirks):
    pool = np.zeros((4096, 2), np.uint32)
    n = C.c_uint32()
    assert O.lib().orc_svo_build(O.vptr(frags), len(frags), level, quirks, O.vptr(pool), len(pool), C.byref(n)) == 0
    return pool, n.value


@pytest.mark.parametrize("mip_test", [True, False])
def test_leaf_average_and_mip_levels_on_a_hand_built_tree(mip_test):
    L = 4  # 16^3 leaves, node resolutions 8 (root children), 4, 2, 1 (leaves)
    block = [(x, y, z) for z in (0, 1) for y in (0, 1) for x in (0, 1)]
    xyz = block + [(15, 15, 15), (3, 5, 7), (3, 5, 7)]
    cols = [0xFF50A0F0] * 8 + [0xFF193264, 0xFF0028C8, 0xFF1E5064]
    frags = _pack(xyz, cols)
    quirks = MIP_TEST if mip_test else 0
    pool, count = _build(frags, L, quirks)
    # leaves: one fragment stores its colour with the alpha nibble (15)
    for c in block:
        assert _rgba(pool[_node(pool, c, 4, L), 1]) == [240, 160, 80, 15]
    assert _rgba(pool[_node(pool, (15, 15, 15), 4, L), 1]) == [100, 50, 25, 15]
    # two fragments in one leaf: (200*15 + 100) / 30 = 103.3, (40*15 + 80) / 30 = 22.7, 30 / 30 = 1, count 30
    assert _rgba(pool[_node(pool, (3, 5, 7), 4, L), 1]) == [103, 22, 1, 30]
    # resolution-2 parents: sum of the 8 children / 8, truncated
    assert _rgba(pool[_node(pool, (0, 0, 0), 3, L), 1]) == [240, 160, 80, 15]
    assert _rgba(pool[_node(pool, (15, 15, 15), 3, L), 1]) == [12, 6, 3, 1]
    # resolution 4: written at mip step 3 (intent) or 4 (the shader's level test)
    assert _rgba(pool[_node(pool, (0, 0, 0), 2, L), 1]) == [30, 20, 10, 1]
    assert _rgba(pool[_node(pool, (15, 15, 15), 2, L), 1]) == [1, 0, 0, 0]
    # resolution 8: only the intent writes it (the level test never matches 8 when L = 4)
    want8 = [0, 0, 0, 0] if mip_test else [3, 2, 1, 0]
    assert _rgba(pool[_node(pool, (0, 0, 0), 1, L), 1]) == want8
    # topology: (3,5,7) shares root child 0 with the block, so 2 root children,
    # 3 depth-2 and 3 depth-3 nodes are subdivided -> 8 + 8 * 8 nodes
    assert count == 8 + 8 * 8


def test_half_coordinates_move_every_fragment_to_the_lower_half():
    L = 3
    frags = _pack([(7, 7, 7), (6, 0, 1)], [0xFF0000FF, 0xFF00FF00])
    pool, _ = _build(frags, L, HALF)
    # (7,7,7) >> 1 = (3,3,3) and (6,0,1) >> 1 = (3,0,0): both in root child 0 (all coordinates < 4)
    assert int(pool[0, 0]) & 0x80000000
    for slot in range(1, 8):
        assert int(pool[slot, 0]) == 0


def _svo_params(dim):
    fields = dict(volume_center=[0.5, 0.5, 0.5], rendering_mode=7, voxel_size=1.0 / dim, volume_dimension=float(dim),
                  trace_start_offset=1.0, indirect_diffuse_intensity=8.0, ambient_occlusion_factor=0.5,
                  min_trace_step_factor=1.0, indirect_specular_intensity=3.0, occlusion_decay=2.0, enable_32_cones=0)
    return O.trace_params(fields, np.eye(4), [0.5, 0.5, -2.0], 3)


def _sample(pool, pos, level, dim=4):
    out = np.zeros(4, np.float32)
    p = _svo_params(dim)
    bmin = (C.c_float * 3)(0, 0, 0)
    bmax = (C.c_float * 3)(1, 1, 1)
    _lib().orc_sample_svo(pool.ctypes.data, C.addressof(p), C.addressof(bmin), C.addressof(bmax),
                          (C.c_float * 3)(*pos), level, out.ctypes.data)
    return [int(round(v * 255)) for v in out]


def test_sample_svo_on_a_hand_built_pool():
    A, B, Cc = 0x80402010, 0xFF00FF00, 0x11223344
    pool = np.zeros((16, 2), np.uint32)
    pool[3] = (0x80000000 | 8, A)  # octant x=1, y=0, z=1 -> slot 1 | 1 << 1 = 3, children at 8
    pool[13, 1] = B                # its sub-cell (0, 1, 1) -> slot 1 | 1 << 2 = 5
    pool[0, 1] = Cc                # octant 0: no children
    at = lambda x, y, z: ((x + 0.5) / 4, (y + 0.5) / 4, (z + 0.5) / 4)  # noqa: E731
    assert _sample(pool, at(2, 1, 3), 0) == _rgba(B)
    assert _sample(pool, at(2, 1, 3), 1) == _rgba(A)
    assert _sample(pool, at(3, 0, 2), 0) == [0, 0, 0, 0]
    # the walk stops at a node without children and returns its colour at every level
    assert _sample(pool, at(0, 0, 0), 0) == _rgba(Cc)
    assert _sample(pool, at(1, 1, 1), 1) == _rgba(Cc)


def test_octree_trace_empty_pool_ambient_occlusion():
    W, H = 4, 3
    half = lambda v: O.lib().orc_float_to_half(v)  # noqa: E731
    g = dict(diffuse=np.zeros((H, W), np.uint32), specular=np.zeros((H, W), np.uint32),
             normal=np.zeros((H, W, 4), np.uint16), emission=np.zeros((H, W, 4), np.uint16),
             depth=np.full((H, W), 0.5, np.float32))
    g["normal"][:] = [half(0.5), half(0.5), half(1.0), half(1.0)]  # n = (0, 0, 1)
    pool = np.zeros((8, 2), np.uint32)
    p = _svo_params(64)
    d = np.zeros((H, W, 4), np.float32)
    s = np.zeros_like(d)
    bmin = (C.c_float * 3)(0, 0, 0)
    bmax = (C.c_float * 3)(1, 1, 1)
    sm_depth = np.ones((4, 4), np.float32)
    sm = O.Shadow(sm_depth.ctypes.data, 4, 4)
    light = dict(direction=[0.0, -1.0, 0.0], intensity=1.0, color=[1.0, 1.0, 1.0], view=np.eye(4), proj=np.eye(4),
                 z_near=0.1, z_far=30.0)
    O.lib().orc_svo_trace(O.vptr(pool), C.byref(O.gbuf_struct(g)), C.byref(p), bmin, bmax, C.byref(sm),
                          C.byref(O.light_struct(light)), 0x3F, 0, H, O.vptr(d), O.vptr(s), None)
    dirs = np.array(json.loads((ROOT / "tests/golden/reference_constants.json").read_text())
                    ["cone_directions_16"]["dirs"], np.float64)
    S = float(np.sum(np.clip(dirs[:, 2], 0.0, None)))
    want = min(1.0, 0.5 * (1.0 + S) / S)
    assert np.allclose(d[..., :3], want, rtol=2e-6, atol=0)


def _bricks(pool, count, level):
    L = O.lib()
    L.orc_svo_bricks.argtypes = [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p]
    L.orc_svo_bricks.restype = C.c_int
    out = np.zeros((count, 3, 3, 3), np.uint32)  # [node, k, j, i]
    assert L.orc_svo_bricks(pool.ctypes.data, count, level, out.ctypes.data) == 0
    return out


def test_bricks_on_a_hand_built_tree():
    # the 2x2x2 block of leaves at cells 0..1 (colour 240, 160, 80, alpha 15) of a 16^3 tree
    L = 4
    block = [(x, y, z) for z in (0, 1) for y in (0, 1) for x in (0, 1)]
    pool, count = _build(_pack(block, [0xFF50A0F0] * 8), L, 0)
    b = _bricks(pool, count, L)
    leaf = _node(pool, (0, 0, 0), 4, L)
    v = [240, 160, 80, 15]
    assert _rgba(b[leaf, 1, 1, 1]) == v                       # its own cell
    assert _rgba(b[leaf, 1, 1, 2]) == v                       # face shared with filled (1,0,0)
    assert _rgba(b[leaf, 1, 1, 0]) == [120, 80, 40, 7]        # face shared with outside (-1,0,0)
    assert _rgba(b[leaf, 0, 0, 0]) == [30, 20, 10, 1]         # corner: 1 filled cell of 8
    assert _rgba(b[leaf, 2, 2, 2]) == v                       # corner inside the filled block
    assert _rgba(b[leaf, 0, 1, 1]) == [120, 80, 40, 7]        # face z = 0 boundary
    assert _rgba(b[leaf, 1, 0, 2]) == [120, 80, 40, 7]        # edge (z centre): 2 filled cells of 4
    assert _rgba(b[leaf, 2, 0, 2]) == [120, 80, 40, 7]        # corner: 4 filled cells of 8 (y = -1 outside)
    assert _rgba(b[leaf, 1, 0, 1]) == [120, 80, 40, 7]        # face y = 0: (0,-1,0) outside


def test_bricks_of_a_uniform_level_are_uniform_inside():
    # every cell of a 4^3 tree filled with one colour: the bricks of interior
    # leaves are that colour everywhere; a boundary texel averages with the
    # outside (one face: half, truncated)
    L = 2
    cells = [(x, y, z) for z in range(4) for y in range(4) for x in range(4)]
    pool, count = _build(_pack(cells, [0xFF649632] * len(cells)), L, 0)  # (50, 150, 100, 15)
    b = _bricks(pool, count, L)
    v = [50, 150, 100, 15]
    for c in [(1, 1, 1), (2, 1, 2), (1, 2, 2)]:
        n = _node(pool, c, 2, L)
        assert all(_rgba(w) == v for w in b[n].reshape(-1))
    edge = _node(pool, (0, 1, 1), 2, L)
    assert _rgba(b[edge, 1, 1, 0]) == [25, 75, 50, 7]
    assert _rgba(b[edge, 1, 1, 1]) == v
