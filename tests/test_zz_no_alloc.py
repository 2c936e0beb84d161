"""vct_trace_noalloc (the allocation-free trace) against vct_trace: the same
kernels on the same inputs, so the diffuse and specular images and the
executed sample counts must be identical, on two calls in a row (the second
on a smaller image, after the static work list held the first call's
entries), in CombinedGI mode (the mode that uses the work list). An image
past the static work list's 2048 tiles is refused with VCT_ERR_CAPACITY
before any launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CAPACITY = 4


def _frame():
    import torch
    import vk_voxel_cone_tracing_amd as V
    from vk_voxel_cone_tracing_amd import passes, scenes

    scene = scenes.room(3, scale=4.0, n_boxes=16, n_spheres=3, emissive_boxes=1)
    ds = V.DeviceScene(scene)
    ctx = V.Context(resolution=32, levels=3, radiance_format=V.FORMAT_RGBA16F, quirks=V.QUIRK_ALL,
                    downsample_region_size=2, extent_l0=4.0)
    ctx.update_clip_regions([2.1, 1.9, 2.05])
    light = scenes.light(3, scale=4.0)
    ld, ls = V.light_structs(light)
    depth = torch.ones((256, 256), dtype=torch.float32, device="cuda")
    V._abi.call("vct_raycast_shadow", ctx._h, None, C.byref(ds.mesh), C.byref(ls), C.c_void_p(depth.data_ptr()),
                256, 256)
    sm = V.shadow_struct(depth)
    mgr = passes.RenderPassManager()
    mgr.add_render_pass("VoxelizationPass", passes.VoxelizationPass(ctx, ds))
    mgr.add_render_pass("RadianceInjectionPass", passes.RadianceInjectionPass(ctx, ds, sm, ld, ls))
    f = V.Frame()
    mgr.draw_single_render_pass("VoxelizationPass", f)
    mgr.draw_single_render_pass("RadianceInjectionPass", f)
    torch.cuda.synchronize()
    return scene, ds, ctx, sm, ld, ls


def _gbuffer(ctx, ds, scene, W, H, seed):
    import torch
    from vk_voxel_cone_tracing_amd import scenes, workloads

    vp, vp_inv, eye = scenes.camera(seed, W, H, scale=4.0)
    g = workloads.gbuffer(ctx._h, ds, scene, (vp, vp_inv, eye), W, H)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    rough = (torch.rand((H, W), generator=gen, device="cuda") * 0.7 + 0.1) * 255.0
    g["diffuse"] = (g["diffuse"] & 0x00FFFFFF) | (torch.round(rough).to(torch.int32) << 24)
    return g, (vp, vp_inv, eye)


def _trace(fn, ctx, g, cam3, sm, ld, ls):
    import torch
    import vk_voxel_cone_tracing_amd as V

    H, W = g["depth"].shape
    mn, ext, vs = ctx.regions[0]
    center = [float(np.float32(np.float32(mn[k]) * np.float32(vs) + np.float32(ext[k]) * np.float32(vs) * 0.5))
              for k in range(3)]
    desc = V._abi.ConeTracingDesc(V._abi.f3(*center), 8, vs, float(ctx.resolution), 1.0, 8.0, 0.5, 1.0, 3.0, 2.0, 0)
    dout = torch.full((H, W, 4), -1.0, dtype=torch.float32, device="cuda")
    sout = torch.full((H, W, 4), -1.0, dtype=torch.float32, device="cuda")
    samples = torch.zeros(1, dtype=torch.int64, device="cuda")
    status = getattr(V._abi.lib(), fn)(ctx._h, None, C.byref(V.gbuffer_struct(g)), C.byref(V.camera_ubo(*cam3)),
                                       C.byref(desc), C.byref(sm), C.byref(ld), C.byref(ls),
                                       C.c_void_p(dout.data_ptr()), C.c_void_p(sout.data_ptr()),
                                       C.c_void_p(samples.data_ptr()))
    torch.cuda.synchronize()
    return status, dout, sout, int(samples.item())


def test_noalloc_trace_equals_trace_on_two_calls():
    import torch

    scene, ds, ctx, sm, ld, ls = _frame()
    try:
        for W, H, seed in ((320, 200, 11), (136, 72, 12)):  # the second call is smaller
            print(f"[no-alloc] {W}x{H}", flush=True)
            g, cam3 = _gbuffer(ctx, ds, scene, W, H, seed)
            st0, d0, s0, n0 = _trace("vct_trace", ctx, g, cam3, sm, ld, ls)
            st1, d1, s1, n1 = _trace("vct_trace_noalloc", ctx, g, cam3, sm, ld, ls)
            assert st0 == 0 and st1 == 0
            assert torch.equal(d0, d1) and torch.equal(s0, s1)
            assert n0 == n1 and n0 > 0
            assert float((s1[..., :3] != 0).float().mean()) > 0.1  # the specular pass ran
    finally:
        ctx.close()


def test_noalloc_trace_refuses_images_past_its_work_list():
    import torch

    scene, ds, ctx, sm, ld, ls = _frame()
    try:
        W, H = 4096, 4096  # 64 x 64 tiles of 64x64 pixels > 2048
        g = dict(diffuse=torch.zeros((H, W), dtype=torch.int32, device="cuda"),
                 normal=torch.zeros((H, W, 4), dtype=torch.int16, device="cuda"),
                 specular=torch.zeros((H, W), dtype=torch.int32, device="cuda"),
                 emission=torch.zeros((H, W, 4), dtype=torch.int16, device="cuda"),
                 depth=torch.ones((H, W), dtype=torch.float32, device="cuda"))
        from vk_voxel_cone_tracing_amd import scenes

        st, _, _, _ = _trace("vct_trace_noalloc", ctx, g, scenes.camera(1, W, H, scale=4.0), sm, ld, ls)
        assert st == CAPACITY
    finally:
        ctx.close()
