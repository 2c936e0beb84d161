"""Pins the oracle's triangle coverage rule, the opacity voxelizer, the
shadow-map visibility and the radiance injection with known answers,
independent brute force (numpy float64 barycentrics) and invariants."""
import ctypes as C

import numpy as np

import oracle_api as O

BIG = 1 << 20


def _list(tris, vs, samples=1, lo=(-BIG,) * 3, hi=(BIG,) * 3):
    tris = np.ascontiguousarray(np.asarray(tris, np.float32).reshape(-1, 9))
    cap = 1 << 20
    out = np.zeros((cap, 4), np.float32)
    n = O.lib().orc_raster_list(O.fptr(tris), len(tris), C.c_float(vs), (C.c_int * 3)(*lo), (C.c_int * 3)(*hi),
                                samples, O.fptr(out), cap)
    return out[:n]


def test_dominant_axis_and_ties():
    L = O.lib()

    def axis(p):
        return L.orc_dominant_axis(O.fptr(np.asarray(p, np.float32)))

    assert axis([0, 0, 0, 1, 0, 0, 0, 1, 0]) == 2  # z-facing
    assert axis([0, 0, 0, 0, 0, 1, 1, 0, 0]) == 1  # y-facing
    assert axis([0, 0, 0, 0, 1, 0, 0, 0, 1]) == 0  # x-facing
    # |n| = (1, 1, 0): x is not strictly greatest, y beats z (msaaVoxelizer.geom:30-31)
    assert axis([0, 0, 0, 0, 0, 1, 1, -1, 0]) == 1
    # |n| = (1, 1, 1): z
    assert axis([1, 0, 0, 0, 1, 0, 0, 0, 1]) == 2


def test_axis_aligned_square_known_answer():
    vs = 0.1
    z = 0.4321
    x0, x1, y0, y1 = 0.23, 0.77, 0.31, 0.89
    quad = [[x0, y0, z, x1, y0, z, x1, y1, z], [x0, y0, z, x1, y1, z, x0, y1, z]]
    s = _list(quad, vs)
    xs = sorted({round(float(v), 6) for v in s[:, 0]})
    ys = sorted({round(float(v), 6) for v in s[:, 1]})
    assert len(s) == 36
    assert xs == [round(0.1 * (i + 0.5), 6) for i in range(2, 8)]
    assert ys == [round(0.1 * (i + 0.5), 6) for i in range(3, 9)]
    assert np.all(s[:, 2] == np.float32(z))
    # 4 samples per cell: the standard 4x positions inside each covered cell
    s4 = _list(quad, vs, samples=4)
    offs = [(0.375, 0.125), (0.875, 0.375), (0.125, 0.625), (0.625, 0.875)]
    want = sum(1 for i in range(12) for j in range(12) for sx, sy in offs
               if x0 < (i + sx) * vs < x1 and y0 < (j + sy) * vs < y1)
    assert len(s4) == want and want > 4 * 30


def _bary_inside(tri2d, q):
    a, b, c = tri2d
    m = np.array([[b[0] - a[0], c[0] - a[0]], [b[1] - a[1], c[1] - a[1]]])
    l1, l2 = np.linalg.solve(m, q - a)
    return 1 - l1 - l2, l1, l2


def test_random_triangles_against_float64_barycentrics():
    rng = np.random.default_rng(12)
    vs = 0.05
    for _ in range(60):
        tri = rng.uniform(-0.3, 0.9, (3, 3)).astype(np.float32)
        n = np.abs(np.cross(tri[1].astype(np.float64) - tri[0], tri[2].astype(np.float64) - tri[0]))
        ax = 0 if (n[0] > n[1] and n[0] > n[2]) else (1 if n[1] > n[2] else 2)
        u, v = [k for k in range(3) if k != ax]
        got = _list(tri.reshape(1, 9), vs)
        got_set = {(int(round(p[u] / vs - 0.5)), int(round(p[v] / vs - 0.5))) for p in got}
        lo = np.floor(tri[:, [u, v]].min(0) / vs).astype(int) - 1
        hi = np.floor(tri[:, [u, v]].max(0) / vs).astype(int) + 1
        t2 = tri[:, [u, v]].astype(np.float64)
        for cu in range(lo[0], hi[0] + 1):
            for cv in range(lo[1], hi[1] + 1):
                q = np.array([np.float32((cu + 0.5) * np.float32(vs)), np.float32((cv + 0.5) * np.float32(vs))],
                             np.float64)
                lam = _bary_inside(t2, q)
                if min(abs(l) for l in lam) < 1e-9:
                    continue  # tie decided by the top-left rule
                inside = min(lam) > 0
                assert inside == ((cu, cv) in got_set), (tri, cu, cv, lam)
        # interpolated coordinate along the dominant axis lies on the triangle's plane
        nn = np.cross(tri[1].astype(np.float64) - tri[0], tri[2].astype(np.float64) - tri[0])
        for p in got[:, :3].astype(np.float64):
            plane = np.dot(nn, p - tri[0]) / np.linalg.norm(nn)
            assert abs(plane) < 1e-5


def test_shared_edges_are_watertight():
    # A fan of triangles whose vertices sit on sample positions: every sample of
    # the polygon is covered exactly once, and two triangulations agree.
    vs = 0.1
    c = (0.55, 0.45)
    ring = [(0.05, 0.05), (0.95, 0.15), (0.85, 0.95), (0.25, 0.85), (0.05, 0.55)]
    fan = []
    for i in range(len(ring)):
        a, b = ring[i], ring[(i + 1) % len(ring)]
        fan.append([c[0], c[1], 0.5, a[0], a[1], 0.5, b[0], b[1], 0.5])
    strip = []
    for i in range(1, len(ring) - 1):
        a, b, d = ring[0], ring[i], ring[i + 1]
        strip.append([a[0], a[1], 0.5, b[0], b[1], 0.5, d[0], d[1], 0.5])
    s_fan = _list(fan, vs)
    s_strip = _list(strip, vs)
    key = lambda s: sorted((round(float(p[0]), 5), round(float(p[1]), 5)) for p in s)
    k_fan, k_strip = key(s_fan), key(s_strip)
    assert len(k_fan) == len(set(k_fan)), "a sample was covered twice"
    assert k_fan == k_strip


def _vox(tris, res=16, vs=0.0625, quirks=0x3F, samples=1):
    vol = O.new_volume(res, 1)
    d = O.vox_desc(((0, 0, 0), (res, res, res), vs), 0, res)
    t = np.ascontiguousarray(np.asarray(tris, np.float32).reshape(-1, 9))
    n = np.zeros_like(t)
    m = np.zeros(len(t), np.uint32)
    O.lib().orc_voxelize_opacity(C.byref(O.vol_struct(vol, res, 1)), C.byref(O.tris_struct(t, n, m)), C.byref(d),
                                 samples, quirks)
    return vol


def test_opacity_voxelizer_square_writes_six_faces():
    res, vs = 16, 0.0625
    z = 5.5 * vs  # through voxel centres of slice 5
    quad = [[0.1, 0.2, z, 0.8, 0.2, z, 0.8, 0.7, z], [0.1, 0.2, z, 0.8, 0.7, z, 0.1, 0.7, z]]
    vol = _vox(quad, res, vs, quirks=0)
    D = res + 2
    want = np.zeros((D, D, D), bool)
    xs = [i for i in range(res) if 0.1 < (i + 0.5) * vs < 0.8]
    ys = [j for j in range(res) if 0.2 < (j + 0.5) * vs < 0.7]
    for i in xs:
        for j in ys:
            want[5 + 1, j + 1, i + 1] = True
    for f in range(6):
        blk = vol[:, :, f * D:(f + 1) * D]
        np.testing.assert_array_equal(blk != 0, want)
        assert np.all(blk[want] == 0xFFFFFFFF)


def test_clamp_x_range_quirk_moves_fragments():
    # With a region whose y range differs from its x range, the reference clamps
    # y into the x range (msaaVoxelizer.frag:46). A square near the region's
    # upper y edge is pulled down to the x range's upper bound.
    res, vs = 16, 0.0625
    D = res + 2
    vol_q, vol_i = O.new_volume(res, 1), O.new_volume(res, 1)
    region = ((0, 4, 0), (res, res, res), vs)  # y spans [0.25, 1.25), x spans [0, 1)
    d = O.vox_desc(region, 0, res)
    y = 1.1  # inside the region, above the x range's max (1.0)
    quad = np.array([[0.3, y, 0.3, 0.6, y, 0.3, 0.6, y, 0.6], [0.3, y, 0.3, 0.6, y, 0.6, 0.3, y, 0.6]], np.float32)
    n = np.zeros_like(quad)
    m = np.zeros(2, np.uint32)
    L = O.lib()
    L.orc_voxelize_opacity(C.byref(O.vol_struct(vol_q, res, 1)), C.byref(O.tris_struct(quad, n, m)), C.byref(d), 1, 0x3F)
    L.orc_voxelize_opacity(C.byref(O.vol_struct(vol_i, res, 1)), C.byref(O.tris_struct(quad, n, m)), C.byref(d), 1, 0x00)
    yq = np.unique(np.argwhere(vol_q[:, :, :D] != 0)[:, 1])
    yi = np.unique(np.argwhere(vol_i[:, :, :D] != 0)[:, 1])
    # intent: fract(1.1 / 1.0) * 16 = 1.6 -> 1; quirk: clamp to 1.0 - vs/4 -> fract(0.984) * 16 = 15
    assert yi.tolist() == [1 + 1] and yq.tolist() == [15 + 1]


def _flat_shadow(value, size=64):
    return np.full((size, size), value, np.float32)


def _down_light(scale=1.0):
    from vk_voxel_cone_tracing_amd import scenes

    L = scenes.light(0, scale)
    d = np.array([0.0, -1.0, 0.0])
    from vk_voxel_cone_tracing_amd import glm

    origin = np.array([0.5, 2.0, 0.5]) * scale
    L["direction"] = d.astype(np.float32)
    L["view"] = glm.look_at_lh(origin, origin + d, np.array([0.0, 0.0, 1.0]))
    L["proj"] = glm.ortho_lh_zo(-scale, scale, -scale, scale, 0.1, 3.0 * scale)
    L["z_near"] = np.float32(0.1)
    L["z_far"] = np.float32(3.0 * scale)
    return L


def test_pcf_visibility_constant_map():
    L = _down_light()
    ol = O.light_struct(L)
    depth = _flat_shadow(0.6)
    sm = O.Shadow(depth.ctypes.data, 64, 64)
    for p in ([0.5, 0.2, 0.5], [0.3, 0.9, 0.7]):
        v = O.lib().orc_visibility(C.byref(sm), C.byref(ol), (C.c_float * 3)(*p), 0x3F)
        assert abs(v - 0.6) < 1e-6  # raw-depth PCF returns the stored depth
    # compare mode: lit when the receiver's NDC depth is not behind the map
    light_z = lambda y: (2.0 - y - 0.1) / (3.0 - 0.1)
    depth2 = _flat_shadow(light_z(0.5))
    sm2 = O.Shadow(depth2.ctypes.data, 64, 64)
    above = O.lib().orc_visibility(C.byref(sm2), C.byref(ol), (C.c_float * 3)(0.5, 0.7, 0.5), 0x00)
    below = O.lib().orc_visibility(C.byref(sm2), C.byref(ol), (C.c_float * 3)(0.5, 0.3, 0.5), 0x00)
    assert above == 1.0 and below == 0.0


def _inject(tris, normals, mats, light, depth, res=16, vs=0.0625, quirks=0x3F):
    vol = O.new_volume(res, 1)
    d = O.vox_desc(((0, 0, 0), (res, res, res), vs), 0, res)
    t = np.ascontiguousarray(np.asarray(tris, np.float32).reshape(-1, 9))
    n = np.ascontiguousarray(np.asarray(normals, np.float32).reshape(-1, 9))
    m = np.zeros(len(t), np.uint32)
    sm = O.Shadow(depth.ctypes.data, depth.shape[1], depth.shape[0])
    O.lib().orc_inject_radiance(C.byref(O.vol_struct(vol, res, 1)), C.byref(O.tris_struct(t, n, m)), mats,
                                C.byref(sm), C.byref(O.light_struct(light)), C.byref(d), 1, quirks)
    return vol


def test_injection_single_lit_triangle_known_answer():
    res, vs = 16, 0.0625
    D = res + 2
    y = 7.5 * vs
    tri = [0.2, y, 0.2, 0.7, y, 0.2, 0.2, y, 0.7]
    nrm = [0, 1, 0] * 3
    mats = (O.Material * 1)()
    albedo = (0.8, 0.4, 0.2, 1.0)
    mats[0].base_color = (C.c_float * 4)(*albedo)
    mats[0].emissive = (C.c_float * 3)(0, 0, 0)
    vol = _inject(tri, nrm, mats, _down_light(), _flat_shadow(1.0))
    # light straight down onto an up-facing triangle: N.L = 1, visibility = 1,
    # radiance = albedo; the -y face (3) receives it with weight |n.y| = 1
    face3 = vol[:, :, 3 * D:4 * D]
    occ = np.argwhere(face3 != 0)
    assert len(occ) > 10 and np.all(occ[:, 1] == 7 + 1)
    want = int(0.8 * 255) | (int(0.4 * 255) << 8) | (int(0.2 * 255) << 16) | (1 << 24)
    assert np.all(face3[face3 != 0] == want)
    # faces 1 and 5 get zero colour with weight 0 but still count the fragment
    for f in (1, 5):
        blk = vol[:, :, f * D:(f + 1) * D]
        assert np.all(blk[face3 != 0] == (1 << 24))
    for f in (0, 2, 4):
        assert not vol[:, :, f * D:(f + 1) * D].any()


def test_injection_emissive_writes_six_faces():
    res = 16
    D = res + 2
    y = 7.5 * 0.0625
    tri = [0.2, y, 0.2, 0.7, y, 0.2, 0.2, y, 0.7]
    mats = (O.Material * 1)()
    mats[0].base_color = (C.c_float * 4)(0.5, 0.5, 0.5, 1.0)
    mats[0].emissive = (C.c_float * 3)(2.0, 0.5, 0.25)  # clamped to [0, 1]
    vol = _inject(tri, [0, 1, 0] * 3, mats, _down_light(), _flat_shadow(0.0))
    want = 255 | (int(0.5 * 255) << 8) | (int(0.25 * 255) << 16) | (1 << 24)
    masks = [vol[:, :, f * D:(f + 1) * D] != 0 for f in range(6)]
    for f in range(6):
        np.testing.assert_array_equal(masks[f], masks[0])
        assert np.all(vol[:, :, f * D:(f + 1) * D][masks[f]] == want)
    assert masks[0].sum() > 10


def test_injection_unlit_fragments_are_discarded():
    # raw-depth visibility of an all-zero map is 0: every lit fragment is discarded
    y = 7.5 * 0.0625
    mats = (O.Material * 1)()
    mats[0].base_color = (C.c_float * 4)(0.5, 0.5, 0.5, 1.0)
    vol = _inject([0.2, y, 0.2, 0.7, y, 0.2, 0.2, y, 0.7], [0, 1, 0] * 3, mats, _down_light(), _flat_shadow(0.0))
    assert not vol.any()
