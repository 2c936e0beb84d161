"""Pins the oracle's clipmap maintenance kernels (clipmapCleaning,
copyAlphaImage, borderWrapping, opacity/radianceDownSample) with closed forms,
brute-force index checks and invariants."""
import ctypes as C

import numpy as np
import pytest

import oracle_api as O

RES, LEVELS = 16, 3
D = RES + 2


def _vol():
    return O.new_volume(RES, LEVELS)


def _view(vol, level, face):
    """(z, y, x) view of one (level, face) block of the reference atlas."""
    return vol[:, level * D:(level + 1) * D, face * D:(face + 1) * D]


def _bytes(t):
    return np.stack([(t >> (8 * k)) & 0xFF for k in range(4)], -1).astype(np.int64)


def _ds(vol, level, prev_min, region=2, radiance=False):
    fn = O.lib().orc_downsample_radiance if radiance else O.lib().orc_downsample_opacity
    fn(C.byref(O.vol_struct(vol, RES, LEVELS)), (C.c_int * 3)(*prev_min), level, region)


def _lerp_closed_form(prev_min, region):
    half = RES // 2
    t = np.zeros((half, half, half))
    for gz in range(half):
        for gy in range(half):
            for gx in range(half):
                g = (gx, gy, gz)
                d = [abs((prev_min[k] >> 1) + g[k] + 0.5 - ((prev_min[k] >> 1) + (half >> 1))) - 0.5 for k in range(3)]
                thr = (half >> 1) - region
                if max(d) >= thr:
                    t[gz, gy, gx] = (max(d) - thr + 1.0) / (region + 1.0)
    return t


def test_downsample_transition_band_closed_form():
    # level L-1 empty (alpha 0) and level L's own .r = 1: the result is the band weight t
    prev_min = (-7, 3, -12)
    vol = _vol()
    vol[:, 1 * D:2 * D, :] = 0x000000FF  # .r = 255 on level 1, all faces
    _ds(vol, 1, prev_min, region=2)
    t = _lerp_closed_form(prev_min, 2)
    half = RES // 2
    for f in range(6):
        blk = _bytes(_view(vol, 1, f))
        for gz in range(half):
            for gy in range(half):
                for gx in range(half):
                    cur = [(prev_min[k] >> 1) + (gx, gy, gz)[k] for k in range(3)]
                    x, y, z = [(c % RES) + 1 for c in cur]
                    a = blk[z, y, x]
                    want = int(np.rint(t[gz, gy, gx] * 255.0))
                    assert a[3] == want and a[1] == want and a[0] == 255, (f, gx, gy, gz)
    assert t.max() > 0 and (t == 0).sum() > 0


def test_downsample_opacity_constant_closed_form():
    # constant alpha a below: every face gives a + (1-a)a = 2a - a^2 outside the band
    a_byte = 100
    a = a_byte / 255.0
    vol = _vol()
    vol[:, 0:D, :] = np.uint32(a_byte << 24)
    prev_min = (-8, -8, -8)
    _ds(vol, 1, prev_min, region=2)
    want = int(np.rint((2 * a - a * a) * 255.0))
    t = _lerp_closed_form(prev_min, 2)
    half = RES // 2
    inner = np.argwhere(t == 0)
    assert len(inner) > 0
    for f in range(6):
        blk = _bytes(_view(vol, 1, f))
        for gz, gy, gx in inner:
            cur = [(prev_min[k] >> 1) + (gx, gy, gz)[k] for k in range(3)]
            x, y, z = [(c % RES) + 1 for c in cur]
            assert blk[z, y, x][3] == want


def test_downsample_radiance_front_to_back_per_face():
    # Texels with offset bit k = 0 are red, bit k = 1 green, all opaque: the
    # positive face along axis k composites the low texel in front (red), the
    # negative face the high texel (green).
    for axis in range(3):
        vol = _vol()
        prev_min = (-8, -8, -8)
        for f in range(6):
            blk = _view(vol, 0, f)
            idx = np.indices(blk.shape)  # (z, y, x)
            coord = idx[2 - axis] - 1  # interior coordinate along the axis
            low = (coord % 2) == 0
            blk[low] = np.uint32(0xFF0000FF)   # red, alpha 1
            blk[~low] = np.uint32(0xFF00FF00)  # green, alpha 1
        _ds(vol, 1, prev_min, region=2, radiance=True)
        t = _lerp_closed_form(prev_min, 2)
        gz, gy, gx = np.argwhere(t == 0)[0]
        cur = [(prev_min[k] >> 1) + (gx, gy, gz)[k] for k in range(3)]
        x, y, z = [(c % RES) + 1 for c in cur]
        pos_face = _bytes(_view(vol, 1, 2 * axis))[z, y, x]
        neg_face = _bytes(_view(vol, 1, 2 * axis + 1))[z, y, x]
        assert pos_face.tolist() == [255, 0, 0, 255], (axis, pos_face)
        assert neg_face.tolist() == [0, 255, 0, 255], (axis, neg_face)


def test_downsample_random_against_float64_brute_force():
    rng = np.random.default_rng(5)
    vol = rng.integers(0, 2**32, size=(D, D * LEVELS, D * 6), dtype=np.uint64).astype(np.uint32)
    before = vol.copy()
    prev_min = (-5, 9, -14)
    _ds(vol, 2, prev_min, region=2)
    pairs = [[(0, 1), (2, 3), (4, 5), (6, 7)], [(1, 0), (3, 2), (5, 4), (7, 6)], [(0, 2), (1, 3), (4, 6), (5, 7)],
             [(2, 0), (3, 1), (6, 4), (7, 5)], [(0, 4), (1, 5), (2, 6), (3, 7)], [(4, 0), (5, 1), (6, 2), (7, 3)]]
    t = _lerp_closed_form(prev_min, 2)
    half = RES // 2
    for f in range(6):
        src = _bytes(_view(before, 1, f))[..., 3] / 255.0
        selfr = _bytes(_view(before, 2, f))[..., 0] / 255.0
        got = _bytes(_view(vol, 2, f))
        for gz in range(0, half, 3):
            for gy in range(0, half, 2):
                for gx in range(half):
                    g = (gx, gy, gz)
                    cur = [(prev_min[k] >> 1) + g[k] for k in range(3)]
                    base = [(2 * c) % RES for c in cur]
                    v = [src[base[2] + (i >> 2) + 1, base[1] + ((i >> 1) & 1) + 1, base[0] + (i & 1) + 1] for i in range(8)]
                    down = sum(v[a] + (1 - v[a]) * v[b] for a, b in pairs[f]) * 0.25
                    x, y, z = [(c % RES) + 1 for c in cur]
                    tt = t[gz, gy, gx]
                    want = down * (1 - tt) + selfr[z, y, x] * tt
                    assert abs(got[z, y, x, 3] - want * 255.0) <= 0.51 + 1e-9
                    assert got[z, y, x, 1] == got[z, y, x, 3]


def test_border_wrap_invariants():
    rng = np.random.default_rng(2)
    for skip_far in (1, 0):
        vol = rng.integers(0, 2**32, size=(D, D * LEVELS, D * 6), dtype=np.uint64).astype(np.uint32)
        before = vol.copy()
        O.lib().orc_wrap_border(C.byref(O.vol_struct(vol, RES, LEVELS)), 2, 6, LEVELS, skip_far)
        for lvl in range(LEVELS):
            for f in range(6):
                b0, b1 = _view(before, lvl, f), _view(vol, lvl, f)
                # the interior never changes
                np.testing.assert_array_equal(b1[1:RES + 1, 1:RES + 1, 1:RES + 1], b0[1:RES + 1, 1:RES + 1, 1:RES + 1])
                for (z, y, x) in [(0, 5, 7), (3, 0, 9), (11, 4, 0), (0, 0, 0)]:
                    src = [((c + RES - 1) % RES) + 1 for c in (z, y, x)]
                    assert b1[z, y, x] == b0[src[0], src[1], src[2]]
                far = [(RES + 1, 5, 7), (3, RES + 1, 9), (11, 4, RES + 1)]
                for (z, y, x) in far:
                    src = [((c + RES - 1) % RES) + 1 for c in (z, y, x)]
                    if skip_far:
                        assert b1[z, y, x] == b0[z, y, x]  # never reached (BorderWrapper.cpp:139-140)
                    else:
                        assert b1[z, y, x] == b0[src[0], src[1], src[2]]


# negative corners come from camera moves toward -x/-y/-z: the slab below the
# region's old minimum wraps to the top of the toroidal volume
@pytest.mark.parametrize("mn", [(12, 0, 14), (-3, -16, 5), (-20, 2, -1)])
def test_clear_region_toroidal(mn):
    rng = np.random.default_rng(4)
    vol = rng.integers(1, 2**32, size=(D, D * LEVELS, D * 6), dtype=np.uint64).astype(np.uint32)
    before = vol.copy()
    ext = (7, 3, 5)
    O.lib().orc_clear_region(C.byref(O.vol_struct(vol, RES, LEVELS)), (C.c_int * 3)(*mn), (C.c_int * 3)(*ext), 1, 6)
    cleared = np.zeros((D, D, D), bool)
    for z in range(ext[2]):
        for y in range(ext[1]):
            for x in range(ext[0]):
                cleared[(z + mn[2]) % RES + 1, (y + mn[1]) % RES + 1, (x + mn[0]) % RES + 1] = True
    for lvl in range(LEVELS):
        for f in range(6):
            b0, b1 = _view(before, lvl, f), _view(vol, lvl, f)
            if lvl == 1:
                assert np.all(b1[cleared] == 0)
                np.testing.assert_array_equal(b1[~cleared], b0[~cleared])
            else:
                np.testing.assert_array_equal(b1, b0)


def test_copy_alpha_moves_alpha_only():
    rng = np.random.default_rng(6)
    rad = rng.integers(0, 2**32, size=(D, D * LEVELS, D * 6), dtype=np.uint64).astype(np.uint32)
    opa = rng.integers(0, 2**32, size=(D, D * LEVELS, D * 6), dtype=np.uint64).astype(np.uint32)
    before = rad.copy()
    O.lib().orc_copy_alpha(C.byref(O.vol_struct(rad, RES, LEVELS)), C.byref(O.vol_struct(opa, RES, LEVELS)), 2, 6)
    for lvl in range(LEVELS):
        for f in range(6):
            r0, r1, o = _view(before, lvl, f), _view(rad, lvl, f), _view(opa, lvl, f)
            if lvl != 2:
                np.testing.assert_array_equal(r1, r0)
                continue
            inner = (slice(1, RES + 1),) * 3
            np.testing.assert_array_equal(r1[inner] & 0x00FFFFFF, r0[inner] & 0x00FFFFFF)
            np.testing.assert_array_equal(r1[inner] >> 24, o[inner] >> 24)
            r1b, r0b = r1.copy(), r0.copy()
            r1b[inner] = 0
            r0b[inner] = 0
            np.testing.assert_array_equal(r1b, r0b)
