"""Known answers for the clip regions of the reference's camera-centred
clipmap: Application::updateClipRegionBoundingBox (Application.cpp:116-128)
centres level L on the camera with half-size (16 >> 1) * 2^L, and
VoxelizationPass::calculateChangeDelta (VoxelizationPass.cpp:438-449) moves
the initial corner -res/2 by trunc(deltaW / minChange) * minChange with
minChange = {2,2,2,2,2,1} voxels (VoxelizationPass.h:57). The expected
corners below were worked out by hand from those lines (for example, level
0 at camera x = -0.49: deltaW = -0.49 - 8 + 8 = -0.49, / 0.25 = -1.96,
trunc -> -1, x 2 -> -2, corner -66; a floor would give -68). Host-only
calls, no device needed."""
import ctypes as C

import pytest

from vk_voxel_cone_tracing_amd import _abi as A


def _regions(cam, res=128, levels=6, extent=16.0):
    cfg = A.Config(res, levels, 0, A.QUIRK_ALL, 1, 10, extent, 0)
    out = (A.ClipmapRegion * levels)()
    A.call("vct_compute_clip_regions", C.byref(cfg), (C.c_float * 3)(*cam), out, levels)
    return [(tuple(r.min_corner), tuple(r.extent), r.voxel_size) for r in out]


def test_camera_at_origin_keeps_the_initial_corners():
    regs = _regions((0.0, 0.0, 0.0))
    for lvl, (mc, ext, vs) in enumerate(regs):
        assert mc == (-64, -64, -64)
        assert ext == (128, 128, 128)
        assert vs == 0.125 * (1 << lvl)


def test_positive_and_negative_offsets():
    regs = _regions((1.3, -2.7, 0.26))
    expect = [(-54, -84, -62), (-60, -74, -64), (-62, -68, -64), (-64, -66, -64), (-64, -64, -64), (-64, -64, -64)]
    assert [r[0] for r in regs] == expect


def test_truncation_toward_zero():
    regs = _regions((-0.49, 0.49, -0.74))
    expect = [(-66, -62, -68), (-64, -64, -66), (-64, -64, -64), (-64, -64, -64), (-64, -64, -64), (-64, -64, -64)]
    assert [r[0] for r in regs] == expect


def test_exact_multiples_of_the_minimum_change():
    regs = _regions((0.5, -0.5, 4.0))
    # level 0: 0.5 / 0.25 = 2 -> 4 voxels; -0.5 -> -4; 4.0 / 0.25 = 16 -> 32
    assert regs[0][0] == (-60, -68, -32)
    # level 5 (minimum change 1 voxel of 4 units): 4.0 / 4 = 1 -> 1 voxel
    assert regs[5][0] == (-64, -64, -63)


def test_other_resolution_and_level_count():
    regs = _regions((3.1, 0.0, -100.0), res=64, levels=3)
    assert [r[2] for r in regs] == [0.25, 0.5, 1.0]
    # level 0: deltaW = 3.1 - 8 + 32 * 0.25 = 3.1, / 0.5 = 6.2 -> 6 -> 12 voxels
    # z: -100 - 8 + 8 = -100, / 0.5 = -200 -> -400 voxels
    assert regs[0][0] == (-20, -32, -432)
    # level 2: deltaW = 3.1, minChange 2 -> 1.55 -> 1 -> 2 voxels; z -100 / 2 = -50 -> -100
    assert regs[2][0] == (-30, -32, -132)


def test_bad_arguments():
    cfg = A.Config(100, 6, 0, 0, 1, 10, 16.0, 0)
    out = (A.ClipmapRegion * 6)()
    assert A.lib().vct_compute_clip_regions(C.byref(cfg), (C.c_float * 3)(0, 0, 0), out, 6) == 1
    cfg = A.Config(128, 6, 0, 0, 1, 10, 16.0, 0)
    assert A.lib().vct_compute_clip_regions(C.byref(cfg), (C.c_float * 3)(0, 0, 0), out, 5) == 1
