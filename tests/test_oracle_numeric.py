"""Pins the oracle's numeric helpers (atomic.glsl pack/unpack, UNORM8 and
half conversions, the RGBA8 running average) against independent
implementations (numpy) and closed forms."""
import ctypes as C

import numpy as np

import oracle_api as O


def _unpack(v):
    out = (C.c_float * 4)()
    O.lib().orc_unpack_rgba8(int(v), out)
    return np.array(out)


def _pack(v):
    return O.lib().orc_pack_rgba8((C.c_float * 4)(*[float(x) for x in v]))


def test_rgba8_pack_unpack_round_trip():
    rng = np.random.default_rng(0)
    words = np.concatenate([rng.integers(0, 2**32, 20000, dtype=np.uint64), [0, 0xFFFFFFFF, 0x000000FF, 0xFF000000]])
    for w in words:
        u = _unpack(int(w))
        # R in the low byte (atomic.glsl:4-12)
        assert u[0] == (int(w) & 0xFF) and u[3] == (int(w) >> 24)
        assert _pack(u) == int(w)


def test_rgba8_pack_truncates_and_wraps():
    # uint() truncates toward zero and & 0xff keeps the low byte (atomic.glsl:14-22)
    assert _pack([1.99, 2.5, 255.9, 256.0]) == (1 | (2 << 8) | (255 << 16) | (0 << 24))
    assert _pack([0, 0, 0, 270.0]) == (14 << 24)


def test_half_conversion_matches_numpy_exhaustive():
    L = O.lib()
    halves = np.arange(65536, dtype=np.uint16)
    ref = halves.view(np.float16).astype(np.float32)
    got = np.array([L.orc_half_to_float(int(h)) for h in halves], np.float32)
    finite = np.isfinite(ref)
    np.testing.assert_array_equal(got[finite].view(np.uint32), ref[finite].view(np.uint32))
    assert np.all(np.isnan(got[~np.isfinite(ref) & np.isnan(ref)]))


def test_float_to_half_round_to_nearest_even():
    L = O.lib()
    rng = np.random.default_rng(1)
    vals = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * 10.0,
        (rng.random(5000).astype(np.float32) * 2e-5),  # subnormal range of half
        np.array([65504.0, 65520.0, 1e6, -1e6, 0.0, -0.0, 5.96e-8, 2.98e-8, 1.0 + 2**-11, 1.0 + 3 * 2**-11],
                 np.float32),
    ])
    got = np.array([L.orc_float_to_half(float(v)) for v in vals], np.uint16)
    with np.errstate(over="ignore"):
        ref = vals.astype(np.float16).view(np.uint16)
    np.testing.assert_array_equal(got, ref)


def test_unorm8_store_rounding():
    L = O.lib()

    def store(v):
        return L.orc_float_to_unorm8_rgba((C.c_float * 4)(v, 0, 0, 0)) & 0xFF

    # every byte survives decode (b / 255) then encode
    for b in range(256):
        assert store(np.float32(b) / np.float32(255.0)) == b
    # round to nearest, ties to even; clamps; NaN -> 0
    assert store(0.5) == 128  # 127.5 -> 128 (even)
    assert store(np.float32(126.5 / 255.0)) in (126, 127)
    assert store(-3.0) == 0 and store(7.0) == 255 and store(float("nan")) == 0


def _avg_seq(values, w=1.0):
    L = O.lib()
    t = 0
    for v in values:
        t = L.orc_rgba8_avg(t, (C.c_float * 4)(*(list(v) + [w])))
    return t


def test_rgba8_running_average_closed_forms():
    # identical values keep their value while the count (alpha) grows by w
    t = _avg_seq([(200, 100, 50)] * 7)
    assert _unpack(t).tolist() == [200, 100, 50, 7]
    # two values: floor of the mean
    t = _avg_seq([(255, 0, 10), (0, 255, 11)])
    assert _unpack(t).tolist() == [127, 127, 10, 2]
    # the count wraps modulo 256 (uint() & 0xff)
    t = _avg_seq([(10, 10, 10)] * 256)
    assert _unpack(t)[3] == 0
    # SVO leaf colours carry a 4-bit alpha of 15 (voxelizer.frag:100): the
    # stored colour is denormalised by the stored count but the new colour is
    # added unweighted, so x_{k+1} = floor((x_k * 15k + v) / (15(k+1)))
    v = [30, 60, 90]
    x = list(v)
    for k in range(1, 3):
        x = [(xi * 15 * k + vi) // (15 * (k + 1)) for xi, vi in zip(x, v)]
    t = _avg_seq([tuple(v)] * 3, w=15.0)
    assert _unpack(t).tolist() == x + [45]
