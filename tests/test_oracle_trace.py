"""Pins the oracle's cone tracer (voxelConeTracing.frag) with closed forms:
calcMinLevel, the data-independent step sequence, the empty-volume and
uniform-opacity results, trilinear sampling of a linear ramp with the
direction-selected faces, the glTF microfacet BRDF, and a whole-pixel
ambient-occlusion known answer."""
import ctypes as C
import math

import numpy as np

import oracle_api as O

RES, LEVELS = 16, 4
D = RES + 2
VS = 0.125
DIM = float(RES)


def _params(center=(0.0, 0.0, 0.0), mode=7, cones32=0, step_factor=1.0, decay=2.0):
    fields = dict(volume_center=list(center), rendering_mode=mode, voxel_size=VS, volume_dimension=DIM,
                  trace_start_offset=1.0, indirect_diffuse_intensity=8.0, ambient_occlusion_factor=0.5,
                  min_trace_step_factor=step_factor, indirect_specular_intensity=3.0, occlusion_decay=decay,
                  enable_32_cones=cones32)
    return O.trace_params(fields, np.eye(4), (0, 0, 0), LEVELS)


def _trace_cone(vol, p, start, d, aperture, max_dist=30.0, start_level=0.0, sf=1.0):
    out = (C.c_float * 4)()
    steps = C.c_uint32(0)
    O.lib().orc_trace_cone(C.byref(O.vol_struct(vol, RES, LEVELS)), C.byref(p), (C.c_float * 3)(*start),
                           (C.c_float * 3)(*d), C.c_float(aperture), C.c_float(max_dist), C.c_float(start_level),
                           C.c_float(sf), out, C.byref(steps))
    return np.array(out), steps.value


def _steps_closed_form(aperture, sf, vs, max_dist, occupied_alpha=None):
    """The march's step positions depend only on aperture, step factor and
    voxel size: t_{k+1} = t_k + max(d_k, vs) * sf with d_k = t_k * 2 tan(a/2)."""
    coef = 2.0 * math.tan(aperture / 2.0)
    t, d, ts = 0.0, vs, []
    while t < max_dist:
        ts.append(t)
        t = t + max(d, vs) * sf
        d = t * coef
    return ts


def test_calc_min_level_closed_form():
    O.lib().orc_calc_min_level.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
    p = _params()
    r0 = VS * DIM * 0.5
    cases = {0.3 * r0: 0.0, 0.75 * r0: 0.5, 1.5 * r0: 1.5, 3.0 * r0: 2.5, 0.9 * r0: 0.8}
    for dist, want in cases.items():
        got = O.lib().orc_calc_min_level(C.byref(p), (C.c_float * 3)(dist, 0.0, 0.0))
        assert abs(got - want) < 1e-5, (dist, got, want)


def test_empty_volume_trace_and_step_count():
    vol = O.new_volume(RES, LEVELS)
    p = _params(center=(100.0, 100.0, 100.0))  # far centre: min level clamps to the top level
    for aperture, sf in ((0.872665, 1.0), (0.3, 0.125), (0.628319, 0.5)):
        out, steps = _trace_cone(vol, p, (0.3, 0.2, 0.1), (0.0, 0.6, 0.8), aperture, sf=sf)
        assert out.tolist() == [0.0, 0.0, 0.0, 1.0]
        assert steps == len(_steps_closed_form(aperture, sf, VS, 30.0))


def test_uniform_opacity_closed_form():
    # every texel of every face/level has alpha a: each step samples a, so
    # opacity_k = 1 - (1 - a)^(seg_k / vs_k) and the occlusion recursion is
    # occ += (1 - occ) * opacity_k / (1 + (t_k + vs_k) * decay).
    a_byte = 40
    a = a_byte / 255.0
    vol = O.new_volume(RES, LEVELS)
    vol[:] = np.uint32(a_byte << 24)
    center = (100.0, 100.0, 100.0)
    p = _params(center=center)
    aperture, sf, decay = 0.872665, 1.0, 2.0
    start = np.array([0.3, 0.2, 0.1])
    d = np.array([0.0, 0.6, 0.8])
    out, steps = _trace_cone(vol, p, start, d, aperture, sf=sf)
    coef = 2.0 * math.tan(aperture / 2.0)
    vs0 = VS
    start_pos = start + d * vs0 * 1.0 * 0.5
    t, diam, seg, occ, res_a = 0.0, VS, vs0, 0.0, 0.0
    r0 = VS * DIM * 0.5
    while t < 30.0 and occ < 1.0:
        pos = start_pos + d * t
        min_level = math.ceil(math.log2(np.linalg.norm(np.array(center) - pos) / r0))
        lvl = min(max(0.0, math.log2(diam / VS), min_level), LEVELS - 1)
        vsl = VS * 2.0 ** lvl
        corr = seg / vsl
        op = min(max(1.0 - (1.0 - a) ** corr, 0.0), 1.0)
        res_a += min(max(1.0 - res_a, 0.0), 1.0) * op
        occ += (1.0 - occ) * op / (1.0 + (t + vsl) * decay)
        prev = t
        t += max(diam, VS) * sf
        seg = t - prev
        diam = t * coef
    assert abs(out[3] - (1.0 - occ)) < 2e-5
    assert out[:3].tolist() == [0.0, 0.0, 0.0]


def test_trilinear_ramp_and_face_selection():
    # face 0 (+x) of level 1 holds a ramp along x: byte value = local index;
    # face 1 (-x) holds zeros. Trilinear filtering of a linear ramp is exact:
    # the sample equals fract(p / extent) * res + 0.5 (in bytes).
    vol = O.new_volume(RES, LEVELS)
    level = 1
    x = np.arange(D, dtype=np.uint32)
    blk = vol[:, level * D:(level + 1) * D, 0:D]
    blk[:] = (x << 24)[None, None, :]
    p = _params()
    extent = VS * 2 ** level * DIM
    out = (C.c_float * 4)()
    for px in (0.37, 1.9, -0.83, 3.3):
        pos = (px, 0.21, -0.4)
        O.lib().orc_sample_clipmap(C.byref(O.vol_struct(vol, RES, LEVELS)), C.byref(p), (C.c_float * 3)(*pos), level,
                                   (C.c_float * 3)(1.0, 0.0, 0.0), out)
        q = px / extent
        want = ((q - math.floor(q)) * RES + 0.5) / 255.0
        assert abs(out[3] - want) < 2e-5, (px, out[3], want)
        O.lib().orc_sample_clipmap(C.byref(O.vol_struct(vol, RES, LEVELS)), C.byref(p), (C.c_float * 3)(*pos), level,
                                   (C.c_float * 3)(-1.0, 0.0, 0.0), out)
        assert out[3] == 0.0


def _brdf_ref(n, v, l, dc, sc, rough):
    """glTF-reference microfacet model in float64 (Schlick F, Smith G with the
    alpha-squared form, GGX/Trowbridge-Reitz D), as cited by brdf.glsl."""
    n, v, l = (np.asarray(x, np.float64) for x in (n, v, l))
    h = (l + v) / np.linalg.norm(l + v)
    ndl = np.clip(n @ l, 0.001, 1.0)
    ndv = np.clip(abs(n @ v), 0.001, 1.0)
    ndh = np.clip(n @ h, 0.0, 1.0)
    vdh = np.clip(v @ h, 0.0, 1.0)
    alpha = rough * rough
    r90 = np.clip(max(sc) * 50.0, 0, 1)
    F = np.asarray(sc) + (r90 - np.asarray(sc)) * (1.0 - vdh) ** 5
    G = (2 * ndl / (ndl + math.sqrt(alpha ** 2 + (1 - alpha ** 2) * ndl ** 2))) * \
        (2 * ndv / (ndv + math.sqrt(alpha ** 2 + (1 - alpha ** 2) * ndv ** 2)))
    Dd = alpha ** 2 / (3.141592 * ((ndh * ndh) * (alpha ** 2 - 1) + 1) ** 2)
    return ndl * ((1 - F) * np.asarray(dc) / 3.141592 + F * G * Dd / (4 * ndl * ndv))


def test_brdf_matches_float64_reference():
    rng = np.random.default_rng(9)
    for _ in range(200):
        n = rng.normal(size=3); n /= np.linalg.norm(n)
        v = rng.normal(size=3); v /= np.linalg.norm(v)
        if n @ v < 0:
            v = -v
        l = rng.normal(size=3); l /= np.linalg.norm(l)
        dc, sc = rng.uniform(0, 1, 3), rng.uniform(0, 1, 3)
        rough = rng.uniform(0.05, 1.0)
        out = (C.c_float * 3)()
        f = lambda a: (C.c_float * 3)(*a.tolist())
        O.lib().orc_brdf(f(n), f(v), f(l), f(dc), f(sc), C.c_float(rough), C.c_float(0.5), out)
        want = _brdf_ref(n, v, l, dc, sc, rough)
        np.testing.assert_allclose(np.array(out), want, rtol=2e-4, atol=1e-6)


def _one_pixel_gbuffer(normal, depth=0.5):
    h2f = lambda x: O.lib().orc_float_to_half(float(x))
    g = dict(
        diffuse=np.array([[0x80FFFFFF]], np.uint32),
        normal=np.array([[[h2f(normal[0] * 0.5 + 0.5), h2f(normal[1] * 0.5 + 0.5), h2f(normal[2] * 0.5 + 0.5),
                           h2f(1.0)]]], np.uint16),
        specular=np.zeros((1, 1), np.uint32),
        emission=np.zeros((1, 1, 4), np.uint16),
        depth=np.array([[depth]], np.float32),
    )
    return g


def test_whole_pixel_ambient_occlusion_empty_volume():
    # With an empty volume every cone returns alpha 1, so
    # AO = (1 + sum_{cos >= 0} cos) / 16 * ambientOcclusionFactor (voxelConeTracing.frag:170-202).
    dirs16 = np.array([
        [0.57735, 0.57735, 0.57735], [0.57735, -0.57735, -0.57735], [-0.57735, 0.57735, -0.57735],
        [-0.57735, -0.57735, 0.57735], [-0.903007, -0.182696, -0.388844], [-0.903007, 0.182696, 0.388844],
        [0.903007, -0.182696, 0.388844], [0.903007, 0.182696, -0.388844], [-0.388844, -0.903007, -0.182696],
        [0.388844, -0.903007, 0.182696], [0.388844, 0.903007, -0.182696], [-0.388844, 0.903007, 0.182696],
        [-0.182696, -0.388844, -0.903007], [0.182696, 0.388844, -0.903007], [-0.182696, 0.388844, 0.903007],
        [0.182696, -0.388844, 0.903007]])
    vol = O.new_volume(RES, LEVELS)
    for normal in ([0.0, 1.0, 0.0], [0.6, 0.0, 0.8]):
        g = _one_pixel_gbuffer(normal)
        p = _params(mode=7)
        n = np.array(normal)
        hn = np.array([O.lib().orc_half_to_float(int(h)) for h in g["normal"][0, 0, :3]]) * 2 - 1
        hn /= np.linalg.norm(hn)
        cos = dirs16 @ hn
        want = (1.0 + cos[cos >= 0].sum()) / 16.0 * 0.5
        rd = np.zeros((1, 1, 4), np.float32)
        rs = np.zeros((1, 1, 4), np.float32)
        O.lib().orc_trace(C.byref(O.vol_struct(vol, RES, LEVELS)), C.byref(O.gbuf_struct(g)), C.byref(p), None, None,
                          0x3F, 0, 1, O.vptr(rd), O.vptr(rs), None)
        assert abs(rd[0, 0, 0] - want) < 1e-5 and rd[0, 0, 0] == rd[0, 0, 1] == rd[0, 0, 2]
        assert n is not None


def test_row_list_trace_equals_row_range_trace():
    # trace_rows (orc_trace_rowlist / _f64, OpenMP over a row list) writes the
    # listed rows exactly as orc_trace / orc_trace_f64 over each row, and
    # leaves the other rows zero.
    rng = np.random.default_rng(5)
    H, W = 6, 5
    vol = O.new_volume(RES, LEVELS)
    vol[...] = rng.integers(0, 2**32, vol.shape, dtype=np.uint64).astype(np.uint32) & np.uint32(0x3F3F3F3F)
    nrm = rng.normal(size=(H, W, 3))
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    h2f = np.vectorize(lambda x: O.lib().orc_float_to_half(float(x)))
    g = dict(
        diffuse=rng.integers(0, 2**32, (H, W), dtype=np.uint64).astype(np.uint32),
        normal=np.ascontiguousarray(h2f(np.concatenate([nrm * 0.5 + 0.5, np.ones((H, W, 1))], -1)), np.uint16),
        specular=rng.integers(0, 2**32, (H, W), dtype=np.uint64).astype(np.uint32),
        emission=np.zeros((H, W, 4), np.uint16),
        depth=rng.uniform(0.2, 0.8, (H, W)).astype(np.float32),
    )
    light = dict(direction=[0.0, -1.0, 0.0], intensity=1.0, color=[1.0, 1.0, 1.0], view=np.eye(4), proj=np.eye(4),
                 z_near=0.1, z_far=10.0)
    shadow = np.ones((4, 4), np.float32)
    p = _params(mode=8)
    rows = np.array([4, 1, 3], np.int32)
    for fp64, fn in ((False, O.lib().orc_trace), (True, O.lib().orc_trace_f64)):
        d, s, n = O.trace_rows(vol, RES, LEVELS, g, p, shadow, light, rows, fp64=fp64)
        dt = np.float64 if fp64 else np.float32
        osm = O.Shadow(shadow.ctypes.data, 4, 4)
        ol = O.light_struct(light)
        total = 0
        for r in rows:
            rd, rs = np.zeros((H, W, 4), dt), np.zeros((H, W, 4), dt)
            cnt = C.c_uint64(0)
            fn(C.byref(O.vol_struct(vol, RES, LEVELS)), C.byref(O.gbuf_struct(g)), C.byref(p), C.byref(osm),
               C.byref(ol), 0x3F, int(r), int(r) + 1, O.vptr(rd), O.vptr(rs), C.byref(cnt))
            np.testing.assert_array_equal(d[r], rd[r])
            np.testing.assert_array_equal(s[r], rs[r])
            total += cnt.value
        assert n == total and n > 0
        others = np.setdiff1d(np.arange(H), rows)
        assert not d[others].any() and not s[others].any()
        assert d[rows].any()


def test_alpha_above_one_saturates_like_alpha_one():
    # A filtered alpha can exceed 1 by rounding (FP32 trilinear weights; the
    # reference's fixed-point filter weights sum to 1 exactly). 1 - a is
    # clamped at 0, so such samples are opaque: a volume of half-float
    # alpha 1 + 2^-10 traces exactly as one of alpha 1 (same colour, the same
    # occlusion, the same early exit). Without the clamp pow(negative,
    # non-integer) is NaN, the opacity clamps to 0 after the first step and
    # the march runs on to the maximum distance.
    h2f = lambda x: O.lib().orc_float_to_half(float(x))  # noqa: E731
    vols = []
    for alpha in (1.0, 1.0009765625):
        vol = O.new_volume(RES, LEVELS, O.FMT_RGBA16F)
        vol[...] = np.array([h2f(0.25), h2f(0.5), h2f(0.75), h2f(alpha)], np.uint16)
        vols.append(vol)
    assert O.lib().orc_half_to_float(int(vols[1].reshape(-1)[3])) > 1.0
    p = _params(center=(100.0, 100.0, 100.0))
    start, d = np.array([0.3, 0.2, 0.1]), np.array([0.0, 0.6, 0.8])
    for sf in (1.0, VS):
        outs = [_trace_cone(vol, p, start, d, 0.2, sf=sf) for vol in vols]
        (o1, n1), (o2, n2) = outs
        assert n1 == n2 and o1[3] < 0.05
        np.testing.assert_array_equal(o1, o2)
