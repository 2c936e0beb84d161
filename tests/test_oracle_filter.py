"""Pins the oracle's SpecularFilterPass (orc_filter.c) by closed forms:
the 32-direction blur of a constant image (its weight count follows from the
float loop of filter.glsl:11-20 and its odd normaliser 8*32 - 15), the
bilateral filter of a constant image, and the Uncharted2 curve."""
import numpy as np
import pytest

import oracle_api as O


def _blur_taps():
    two_pi = np.float32(6.28318530718)
    step = np.float32(two_pi / np.float32(32.0))
    d, n = np.float32(0.0), 0
    while d < two_pi:
        n += 1
        d = np.float32(d + step)
    rings, i = 0, np.float32(0.125)
    while i <= np.float32(1.0):
        rings += 1
        i = np.float32(i + np.float32(0.125))
    return n, rings


def test_blur_of_constant_image_closed_form():
    H, W = 12, 17
    rng = np.random.default_rng(0)
    diffuse = rng.uniform(0, 1, (H, W, 4)).astype(np.float32)
    c = np.array([0.3, 0.6, 0.9, 1.0], np.float32)
    spec = np.broadcast_to(c, (H, W, 4)).copy()
    out = O.specular_filter(diffuse, spec, method=1)
    n, rings = _blur_taps()
    assert rings == 8 and n in (32, 33)
    want = diffuse[..., :3] + c[:3] * (1 + n * rings) / 241.0
    np.testing.assert_allclose(out[..., :3], want, rtol=2e-6, atol=2e-6)
    np.testing.assert_array_equal(out[..., 3], diffuse[..., 3])


def test_bilateral_of_constant_image_is_identity():
    H, W = 9, 11
    rng = np.random.default_rng(1)
    diffuse = rng.uniform(0, 1, (H, W, 4)).astype(np.float32)
    c = np.array([0.25, 0.5, 0.75, 1.0], np.float32)
    out = O.specular_filter(diffuse, np.broadcast_to(c, (H, W, 4)).copy(), method=0)
    np.testing.assert_allclose(out[..., :3], diffuse[..., :3] + c[:3], rtol=1e-6, atol=1e-6)


def test_bilateral_keeps_edges_blur_does_not():
    # a step edge: the bilateral filter (range sigma 10 on [0, 1] values is
    # wide) still weights by the spatial kernel only, the blur spreads it
    H, W = 8, 32
    spec = np.zeros((H, W, 4), np.float32)
    spec[:, W // 2:, :3] = 1.0
    zero = np.zeros_like(spec)
    bil = O.specular_filter(zero, spec, method=0)
    assert 0.0 < bil[4, W // 2 - 1, 0] < 0.5 < bil[4, W // 2, 0] < 1.0
    assert abs(bil[4, 0, 0]) < 1e-6 and abs(bil[4, W - 1, 0] - 1.0) < 1e-6


@pytest.mark.parametrize("gamma,exposure", [(2.2, 0.1), (1.0, 1.0), (2.4, 0.5)])
def test_tonemap_closed_form(gamma, exposure):
    H, W = 4, 5
    rng = np.random.default_rng(2)
    diffuse = rng.uniform(0, 4, (H, W, 4)).astype(np.float32)
    spec = np.zeros_like(diffuse)
    spec[..., 3] = 1.0
    out = O.specular_filter(diffuse, spec, gamma, exposure, tonemap=1, method=0)

    def u2(x):
        A, B, C, D, E, F = 0.15, 0.50, 0.10, 0.20, 0.02, 0.30
        return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F

    want = (u2(diffuse[..., :3].astype(np.float64) * exposure) / u2(11.2)) ** (1.0 / gamma)
    np.testing.assert_allclose(out[..., :3], want, rtol=5e-6)
