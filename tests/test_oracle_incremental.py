"""Toroidal incremental revoxelization on the oracle (intent mode, half-open
voxel cells): along a camera path, clearing and voxelizing only the regions a
move exposes - one slab per moved axis (fillRevoxelizationRegions,
VoxelizationPass.cpp:450-494) and, on a level >= 1, the old downsampled inner
box - then rebuilding the mips and borders gives the same opacity atlas as
voxelizing every level of the moved regions from scratch. The room's walls lie
on voxel faces, so region faces cut through geometry."""
import ctypes as C

import numpy as np
import pytest

import oracle_api as O
from vk_voxel_cone_tracing_amd import scenes

RES, LEVELS, VS0, MIN_CHANGE = 32, 3, 0.125, 2
INTENT = 0x3F & ~1  # every quirk but the closed-region x-range clamp


def _regions(cam):
    """Per level, the region's min cell: centred on the camera and snapped to
    the minimum change, so level i-1's region sits inside level i's."""
    out = []
    for i in range(LEVELS):
        vs = VS0 * 2 ** i
        c = np.floor(np.asarray(cam) / vs / MIN_CHANGE).astype(int) * MIN_CHANGE
        out.append(tuple(int(x) - RES // 2 for x in c))
    return out


def _revox_boxes(old, new):
    """The regions to clear and voxelize again after old -> new (the rule of
    vct_move_clip_regions)."""
    out = []
    for i in range(LEVELS):
        d = np.subtract(new[i], old[i])
        if np.any(np.abs(d) >= RES):
            out.append([(new[i], (RES,) * 3)])
            continue
        boxes = []
        for k in range(3):
            if d[k] == 0:
                continue
            mn, ext = list(new[i]), [RES] * 3
            ext[k] = abs(int(d[k]))
            mn[k] = new[i][k] + RES - ext[k] if d[k] > 0 else new[i][k]
            boxes.append((tuple(mn), tuple(ext)))
        if i > 0 and any((a >> 1) != (b >> 1) for a, b in zip(old[i - 1], new[i - 1])):
            lo = [max(o >> 1, n) for o, n in zip(old[i - 1], new[i])]
            hi = [min((o >> 1) + RES // 2, n + RES) for o, n in zip(old[i - 1], new[i])]
            if all(h > l for l, h in zip(lo, hi)):
                boxes.append((tuple(lo), tuple(h - l for l, h in zip(lo, hi))))
        out.append(boxes)
    return out


class _Atlas:
    def __init__(self, tris, samples):
        self.vol = O.new_volume(RES, LEVELS)
        self.tris, self.samples = tris, samples
        self.L = O.lib()

    def _s(self):
        return C.byref(O.vol_struct(self.vol, RES, LEVELS))

    def voxelize(self, level, mn, ext):
        vs = VS0 * 2 ** level
        d = O.vox_desc((mn, ext, vs), level, RES)
        d.clip_max_extent = float(np.float32(RES * vs))
        self.L.orc_voxelize_opacity(self._s(), C.byref(self.tris), C.byref(d), self.samples, INTENT)

    def clear(self, level, mn, ext):
        self.L.orc_clear_region(self._s(), (C.c_int * 3)(*mn), (C.c_int * 3)(*ext), level, 6)

    def finish(self, regions):
        for level in range(1, LEVELS):
            self.L.orc_downsample_opacity(self._s(), (C.c_int * 3)(*regions[level - 1]), level, 2)
        self.L.orc_wrap_border(self._s(), 2, 6, LEVELS, 1)


def _full(tris, samples, regions):
    a = _Atlas(tris, samples)
    for level in range(LEVELS):
        a.voxelize(level, regions[level], (RES,) * 3)
    a.finish(regions)
    return a.vol


@pytest.mark.parametrize("samples", [1, 4])
def test_incremental_equals_full_along_camera_path(samples):
    sc = scenes.room(0, scale=4.0, n_boxes=24, n_spheres=4, emissive_boxes=2)
    pos, nrm, mat = sc.world_triangles()
    pos, nrm = np.ascontiguousarray(pos, np.float32), np.ascontiguousarray(nrm, np.float32)
    mat = np.ascontiguousarray(mat, np.uint32)
    tris = O.tris_struct(pos, nrm, mat)
    base = np.full(3, 2.0)
    cams = [base, base + [0.61, -0.27, 0.93], base + [0.65, 0.9, 0.2], base + [-1.7, 0.95, -0.4],
            base + [-1.1, 1.2, 0.3], base + [30.0, 0.0, 0.0]]
    regions = _regions(cams[0])
    inc = _Atlas(tris, samples)
    for level in range(LEVELS):
        inc.voxelize(level, regions[level], (RES,) * 3)
    inc.finish(regions)
    kinds = set()
    for cam in cams[1:]:
        new = _regions(cam)
        boxes = _revox_boxes(regions, new)
        for level, bl in enumerate(boxes):
            for mn, ext in bl:
                kinds.add((level, "full" if ext == (RES,) * 3 else "part"))
                inc.clear(level, mn, ext)
                inc.voxelize(level, mn, ext)
        regions = new
        inc.finish(regions)
        ref = _full(tris, samples, regions)
        assert np.count_nonzero(ref) or cam is cams[-1]
        np.testing.assert_array_equal(inc.vol, ref, err_msg=f"camera {cam}")
    # every level took partial moves, and the last jump rebuilt every level
    assert {(lv, "part") for lv in range(LEVELS)} <= kinds and (LEVELS - 1, "full") in kinds


def test_closed_region_clamp_breaks_slab_exactness():
    # The reference's closed region puts a wall lying on the region's max face
    # into the last texel; after a move that face is interior and the wall
    # belongs to the next texel, which a slab update alone cannot fix.
    vs = VS0
    quad = np.array([[1.0, 0.1, 0.1, 1.0, 0.9, 0.1, 1.0, 0.9, 0.9],
                     [1.0, 0.1, 0.1, 1.0, 0.9, 0.9, 1.0, 0.1, 0.9]], np.float32)  # x = 1.0 = cell 8
    n = np.tile(np.array([1, 0, 0] * 3, np.float32), (2, 1))
    m = np.zeros(2, np.uint32)
    tris = O.tris_struct(quad, n, m)
    L = O.lib()
    for quirks, last in ((0x3F, True), (INTENT, False)):
        vol = O.new_volume(RES, 1)
        d = O.vox_desc(((-24, 0, 0), (RES, RES, RES), vs), 0, RES)  # x cells [-24, 8): the wall at the max face
        L.orc_voxelize_opacity(C.byref(O.vol_struct(vol, RES, 1)), C.byref(tris), C.byref(d), 1, quirks)
        xs = np.unique(np.argwhere(vol[:, :, :RES + 2] != 0)[:, 2]) - 1
        # parity: clamped into cell 7 (texel 7); intent: cell 8 is outside the region
        assert xs.tolist() == ([7] if last else []), (quirks, xs)
