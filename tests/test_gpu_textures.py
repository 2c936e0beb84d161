"""GPU parity of the scene-texture lookups of the voxel passes against the
oracle, on the textured room (base-colour, emissive and occlusion textures,
walls tiled 3 x 3 so REPEAT addressing is exercised, a 5 x 3 texture for
odd sizes): opacity voxelization (msaaVoxelizer.frag:64), radiance
injection (msaaInjectRadiance.frag:73-82,131-136) and the octree fragment
list and node pool (voxelizer.frag:52-77) must be bit-identical. The
textures must also change the result (the untextured room differs), so the
comparison is not vacuous.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_api as O
from test_gpu_clipmap import EXTENT_L0, LEVELS, RES, SCALE, _desc_to_oracle, _light_and_shadow

pytestmark = pytest.mark.gpu


def _voxel_passes(textured, fmt, quirks, samples):
    import torch
    import vk_voxel_cone_tracing_amd as V
    from vk_voxel_cone_tracing_amd import scenes

    scene = scenes.room(0, scale=SCALE, n_boxes=24, n_spheres=4, emissive_boxes=2, textured=textured)
    ctx = V.Context(resolution=RES, levels=LEVELS, radiance_format=fmt, quirks=quirks, raster_samples=samples,
                    downsample_region_size=2, extent_l0=EXTENT_L0)
    try:
        ctx.update_clip_regions(np.full(3, 0.5 * SCALE) + np.array([0.13, -0.07, 0.21]))
        ds = V.DeviceScene(scene)
        pos, nrm, mat = scene.world_triangles()
        light, depth, depth_d = _light_and_shadow(scene, pos, nrm, mat)
        ld, ls = V.light_structs(light)
        sm = V.shadow_struct(depth_d)
        torch.cuda.synchronize()
        ctx.clear_volume(V.VOLUME_OPACITY)
        ctx.clear_volume(V.VOLUME_RADIANCE)
        ref_o, ref_r = O.new_volume(RES, LEVELS), O.new_volume(RES, LEVELS, fmt)
        L = O.lib()
        tris, mats = O.tris_struct(pos, nrm, mat, scene), O.materials_struct(scene)
        osm = O.Shadow(depth.ctypes.data, depth.shape[1], depth.shape[0])
        olight = O.light_struct(light)
        for level in range(LEVELS):
            desc = ctx.voxelization_desc(level)
            odesc = _desc_to_oracle(desc, ctx.regions[level])
            ctx.voxelize_opacity(ds, desc, level)
            ctx.inject_radiance(ds, sm, ld, ls, desc, level)
            L.orc_voxelize_opacity_textured(C.byref(O.vol_struct(ref_o, RES, LEVELS)), C.byref(tris), mats,
                                            C.byref(odesc), samples, quirks)
            L.orc_inject_radiance(C.byref(O.vol_struct(ref_r, RES, LEVELS)), C.byref(tris), mats, C.byref(osm),
                                  C.byref(olight), C.byref(odesc), samples, quirks)
        return ctx.read_volume(V.VOLUME_OPACITY), ctx.read_volume(V.VOLUME_RADIANCE), ref_o, ref_r
    finally:
        ctx.close()


@pytest.mark.parametrize("fmt,quirks,samples", [(0, 0x3F, 1), (1, 0x3F, 1), (1, 0x00, 4)])
def test_textured_voxel_passes_bit_exact(fmt, quirks, samples):
    print(f"[textures] clipmap fmt={fmt} quirks={quirks:#x} samples={samples}", flush=True)
    got_o, got_r, ref_o, ref_r = _voxel_passes(True, fmt, quirks, samples)
    np.testing.assert_array_equal(got_o, ref_o)
    np.testing.assert_array_equal(got_r, ref_r)
    _, _, plain_o, plain_r = _voxel_passes(False, fmt, quirks, samples)
    n_tex, n_plain = int((ref_o != 0).sum()), int((plain_o != 0).sum())
    changed = int((ref_r != plain_r).any(-1).sum()) if ref_r.ndim == 4 else int((ref_r != plain_r).sum())
    print(f"opacity texels textured {n_tex} / untextured {n_plain}; radiance texels changed {changed}", flush=True)
    assert 0 < n_tex < n_plain and changed > 100


@pytest.mark.parametrize("quirks,level", [(0x3F, 7), (0x00, 7)])
def test_textured_octree_bit_exact(quirks, level):
    import torch
    import vk_voxel_cone_tracing_amd as V
    from vk_voxel_cone_tracing_amd import scenes

    print(f"[textures] octree quirks={quirks:#x} level={level}", flush=True)
    out = {}
    for textured in (True, False):
        scene = scenes.room(1, scale=1.0, n_boxes=24, n_spheres=4, emissive_boxes=2, textured=textured)
        pos, nrm, mat = scene.world_triangles()
        ctx = V.Context(resolution=16, levels=2, quirks=quirks, downsample_region_size=1, extent_l0=1.0)
        try:
            ds = V.DeviceScene(scene)
            light = scenes.light(1, scale=1.0)
            depth = O.raycast_shadow(pos, nrm, mat, light, 128, 128)
            sm = V.shadow_struct(torch.from_numpy(depth).cuda())
            ld, ls = V.light_structs(light)
            nfrag = C.c_uint32(0)
            V._abi.call("vct_svo_voxelize", ctx._h, None, C.byref(ds.mesh), C.byref(sm), C.byref(ld), C.byref(ls),
                        level, C.byref(nfrag))
            n = C.c_size_t(0)
            frags = np.zeros((nfrag.value, 2), np.uint32)
            V._abi.call("vct_svo_read_fragments", ctx._h, None, frags.ctypes.data_as(C.c_void_p), nfrag.value,
                        C.byref(n))
            V._abi.call("vct_svo_build", ctx._h, None)
            V._abi.call("vct_svo_read_nodes", ctx._h, None, None, 0, C.byref(n))
            nodes = np.zeros((n.value, 2), np.uint32)
            V._abi.call("vct_svo_read_nodes", ctx._h, None, nodes.ctypes.data_as(C.c_void_p), n.value, C.byref(n))
        finally:
            ctx.close()
        L = O.lib()
        cap = 4 * max(nfrag.value, 1) + 16
        ofr = np.zeros((cap, 2), np.uint32)
        osm = O.Shadow(depth.ctypes.data, 128, 128)
        onf = L.orc_svo_fragments(C.byref(O.tris_struct(pos, nrm, mat, scene)), O.materials_struct(scene),
                                  C.byref(osm), C.byref(O.light_struct(light)),
                                  (C.c_float * 3)(*scene.bb_min.tolist()), (C.c_float * 3)(*scene.bb_max.tolist()),
                                  1 << level, 1, quirks, O.vptr(ofr), cap)
        opool = np.zeros((max(1000000, 8 * onf), 2), np.uint32)
        ocount = C.c_uint32(0)
        assert L.orc_svo_build(O.vptr(ofr), onf, level, quirks, O.vptr(opool), len(opool), C.byref(ocount)) == 0
        np.testing.assert_array_equal(frags, ofr[:onf])
        np.testing.assert_array_equal(nodes, opool[:ocount.value])
        out[textured] = frags
    print(f"fragments textured {len(out[True])} / untextured {len(out[False])}", flush=True)
    assert 1000 < len(out[True]) < len(out[False])
