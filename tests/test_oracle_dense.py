"""Pins the oracle's dense "plain" method (orc_dense.c): voxelization known
answers with first-triangle-wins colours, the anisotropic mip pyramid's closed
forms, and the empty-grid ambient-occlusion closed form of the march."""
import ctypes as C

import numpy as np

import oracle_api as O


def _tris(tris, mats_rgb):
    t = np.ascontiguousarray(np.asarray(tris, np.float32).reshape(-1, 9))
    n = np.zeros_like(t)
    m = np.arange(len(t), dtype=np.uint32)
    mats = (O.Material * len(t))()
    for i, rgb in enumerate(mats_rgb):
        mats[i].base_color = (C.c_float * 4)(rgb[0], rgb[1], rgb[2], 1.0)
    return t, n, m, mats


def test_voxelize_square_first_triangle_wins():
    res = 16
    vs = 1.0 / res
    z = 4.5 * vs
    quad = [[0.1, 0.2, z, 0.8, 0.2, z, 0.8, 0.7, z], [0.1, 0.2, z, 0.8, 0.7, z, 0.1, 0.7, z],
            [0.1, 0.2, z, 0.8, 0.2, z, 0.8, 0.7, z]]  # a duplicate of triangle 0 drawn last
    t, n, m, mats = _tris(quad, [(1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)])
    d, lv = O.new_dense(res)
    O.lib().orc_dense_voxelize(C.byref(d), C.byref(O.tris_struct(t, n, m)), mats, 1)
    occ = lv[0][0] != 0
    xs = [i for i in range(res) if 0.1 < (i + 0.5) * vs < 0.8]
    ys = [j for j in range(res) if 0.2 < (j + 0.5) * vs < 0.7]
    want = np.zeros_like(occ)
    for i in xs:
        for j in ys:
            want[4 + 1, j + 1, i + 1] = True
    np.testing.assert_array_equal(occ, want)
    vals = set(int(v) for v in lv[0][0][occ])
    # red where triangle 0 covers the sample, green elsewhere; never blue (drawn last)
    assert vals <= {0xFF0000FF, 0xFF00FF00} and 0xFF0000FF in vals and 0xFF00FF00 in vals


def test_mip_constant_closed_form():
    res = 8
    d, lv = O.new_dense(res)
    a_b, c_b = 90, 40
    lv[0][0][1:-1, 1:-1, 1:-1] = np.uint32(c_b | (c_b << 8) | (c_b << 16) | (a_b << 24))
    O.lib().orc_dense_mip(C.byref(d))
    a, c = a_b / 255.0, c_b / 255.0
    want_a = int(np.rint((2 * a - a * a) * 255.0))
    want_c = int(np.rint(c * (2 - a) * 255.0))
    got = lv[1][:, 1:-1, 1:-1, 1:-1]
    assert np.all(((got >> 24) & 0xFF) == want_a)
    assert np.all((got & 0xFF) == want_c)
    # the pyramid reaches 1^3 and keeps a constant grid constant at every level
    assert lv[-1].shape == (6, 3, 3, 3)
    assert len(np.unique(lv[-1][:, 1, 1, 1])) == 1
    # borders stay zero
    assert not lv[1][:, 0].any() and not lv[1][:, -1].any()


def _one_pixel(normal):
    h2f = lambda x: O.lib().orc_float_to_half(float(x))
    return dict(
        diffuse=np.array([[0xFF808080]], np.uint32),
        normal=np.array([[[h2f(normal[0] * 0.5 + 0.5), h2f(normal[1] * 0.5 + 0.5), h2f(normal[2] * 0.5 + 0.5),
                           h2f(1.0)]]], np.uint16),
        specular=np.zeros((1, 1), np.uint32),
        emission=np.zeros((1, 1, 4), np.uint16),
        depth=np.array([[0.5]], np.float32),
    )


def test_empty_grid_ambient_occlusion_closed_form():
    dirs = np.array([
        [0.57735, 0.57735, 0.57735], [0.57735, -0.57735, -0.57735], [-0.57735, 0.57735, -0.57735],
        [-0.57735, -0.57735, 0.57735], [-0.903007, -0.182696, -0.388844], [-0.903007, 0.182696, 0.388844],
        [0.903007, -0.182696, 0.388844], [0.903007, 0.182696, -0.388844], [-0.388844, -0.903007, -0.182696],
        [0.388844, -0.903007, 0.182696], [0.388844, 0.903007, -0.182696], [-0.388844, 0.903007, 0.182696],
        [-0.182696, -0.388844, -0.903007], [0.182696, 0.388844, -0.903007], [-0.182696, 0.388844, 0.903007],
        [0.182696, -0.388844, 0.903007]])
    d, lv = O.new_dense(16)
    g = _one_pixel([0.0, 0.0, 1.0])
    p = O.dense_params(np.eye(4))
    out = np.zeros((1, 1, 4), np.float64)
    O.lib().orc_dense_trace(C.byref(d), C.byref(O.gbuf_struct(g)), C.byref(p), 0, 1, O.vptr(out), None)
    n = np.array([O.lib().orc_half_to_float(int(h)) for h in g["normal"][0, 0, :3]]) * 2 - 1
    n /= np.linalg.norm(n)
    cos = dirs @ n
    assert abs(out[0, 0, 3] - (1 + cos[cos >= 0].sum()) / 16 * 0.5) < 1e-12
    assert out[0, 0, :3].tolist() == [0.0, 0.0, 0.0]


def test_opaque_grid_closed_form():
    # Opaque white grid, pixel at the centre, cones short enough to stay inside:
    # every sample has alpha = rgb = sum(w) = 1 (to 1e-5), so each cone's colour
    # is corr_0 = 1 and its occlusion follows occ += (1 - occ) / (1 + (t_k + vs_k) decay)
    # over the data-independent step sequence.
    res = 16
    d, lv = O.new_dense(res)
    lv[0][0][1:-1, 1:-1, 1:-1] = np.uint32(0xFFFFFFFF)
    O.lib().orc_dense_mip(C.byref(d))
    g = _one_pixel([0.0, 0.0, 1.0])
    T = np.eye(4)
    T[:3, 3] = [0.5, 0.5, 0.0]  # NDC (0, 0, 0.5) -> grid position (0.5, 0.5, 0.5)
    max_dist = 0.2
    p = O.dense_params(T, max_dist=max_dist)
    out = np.zeros((1, 1, 4), np.float64)
    O.lib().orc_dense_trace(C.byref(d), C.byref(O.gbuf_struct(g)), C.byref(p), 0, 1, O.vptr(out), None)
    vs, coef, decay, nl = 1.0 / res, 2 * np.tan(0.872665 / 2), 2.0, 5
    t, diam, occ = 0.0, vs, 0.0
    while t < max_dist and occ < 1.0:
        lvl = min(max(np.log2(diam / vs), 0.0), nl - 1)
        occ += (1 - occ) / (1 + (t + vs * 2 ** lvl) * decay)
        t += max(diam, vs) * 1.0
        diam = t * coef
    n = np.array([O.lib().orc_half_to_float(int(h)) for h in g["normal"][0, 0, :3]]) * 2 - 1
    n /= np.linalg.norm(n)
    dirs = np.array([[0.57735, 0.57735, 0.57735], [0.57735, -0.57735, -0.57735], [-0.57735, 0.57735, -0.57735],
                     [-0.57735, -0.57735, 0.57735], [-0.903007, -0.182696, -0.388844], [-0.903007, 0.182696, 0.388844],
                     [0.903007, -0.182696, 0.388844], [0.903007, 0.182696, -0.388844],
                     [-0.388844, -0.903007, -0.182696], [0.388844, -0.903007, 0.182696],
                     [0.388844, 0.903007, -0.182696], [-0.388844, 0.903007, 0.182696],
                     [-0.182696, -0.388844, -0.903007], [0.182696, 0.388844, -0.903007],
                     [-0.182696, 0.388844, 0.903007], [0.182696, -0.388844, 0.903007]])
    cos = dirs @ n
    cs = cos[cos >= 0].sum()
    assert abs(out[0, 0, 3] - (1 + cs * (1 - occ)) / 16 * 0.5) < 1e-4
    albedo = 0x80 / 255.0
    np.testing.assert_allclose(out[0, 0, :3], cs / 16 * albedo * 8.0, rtol=1e-4)


def test_voxelize_is_translation_invariant():
    # Shifting the triangles and the grid by the same dyadic offset (exact in
    # float) must reproduce the unshifted grid texel for texel.
    res = 16
    rng = np.random.default_rng(9)
    # coordinates on a 2^-16 lattice so that adding the offset is exact
    tris = (np.round(rng.uniform(0.05, 0.95, size=(40, 9)) * 65536.0) / 65536.0).astype(np.float32)
    t, n, m, mats = _tris(tris, [tuple(rng.uniform(0, 1, 3)) for _ in range(40)])
    d0, lv0 = O.new_dense(res)
    O.lib().orc_dense_voxelize(C.byref(d0), C.byref(O.tris_struct(t, n, m)), mats, 4)
    off = np.array([3.5, -2.25, 8.0], np.float32)
    ts = np.ascontiguousarray((t.reshape(-1, 3) + off).reshape(-1, 9))
    assert np.array_equal((ts.reshape(-1, 3) - off).reshape(-1, 9), t)
    d1, lv1 = O.new_dense(res, grid_min=tuple(off.tolist()))
    O.lib().orc_dense_voxelize(C.byref(d1), C.byref(O.tris_struct(ts, n, m)), mats, 4)
    assert (lv0[0] != 0).sum() > 100
    np.testing.assert_array_equal(lv0[0], lv1[0])
