"""The drop-in boundary for scene textures. The clipmap voxel passes and
the octree voxelizer sample base-colour, emissive and occlusion textures
(msaaVoxelizer.frag:64, msaaInjectRadiance.frag:73-82,131-136,
voxelizer.frag:52-77); texture indices outside the mesh's texture table, or
textured materials without texcoords, are invalid arguments. The dense
method, the batched queue and the ray-cast G-buffer use material factors
only and refuse textured materials with VCT_ERR_UNSUPPORTED; the shadow
ray-cast ignores materials. Every refusal happens before any kernel runs."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INVALID, UNSUPPORTED = 1, 5


def _scene(textured=True):
    from vk_voxel_cone_tracing_amd import scenes

    return scenes.room(0, scale=1.0, n_boxes=4, n_spheres=1, textured=textured)


def _passes(ctx, ds, s):
    """Status of each voxel pass on the scene."""
    import torch

    import vk_voxel_cone_tracing_amd as V
    from vk_voxel_cone_tracing_amd import _abi as A
    from vk_voxel_cone_tracing_amd import scenes

    L = A.lib()
    ld, ls = V.light_structs(scenes.light(0, scale=1.0))
    depth = torch.ones((64, 64), dtype=torch.float32, device="cuda")
    sm = V.shadow_struct(depth)
    d = ctx.voxelization_desc(0)
    n = C.c_uint32()
    out = dict(opacity=L.vct_voxelize_opacity(ctx._h, s, C.byref(ds.mesh), C.byref(d), 0),
               inject=L.vct_inject_radiance(ctx._h, s, C.byref(ds.mesh), C.byref(sm), C.byref(ld), C.byref(ls),
                                            C.byref(d), 0),
               svo=L.vct_svo_voxelize(ctx._h, s, C.byref(ds.mesh), C.byref(sm), C.byref(ld), C.byref(ls), 4,
                                      C.byref(n)))
    g = C.c_void_p()
    A.call("vct_dense_create", C.byref(g), C.byref(A.DenseDesc(A.f3(-0.1, -0.1, -0.1), 1.2, 16, 0, 1, 0)))
    try:
        out["dense"] = L.vct_dense_voxelize(g, s, C.byref(ds.mesh))
    finally:
        A.call("vct_dense_destroy", g)
    W = H = 8
    surf = torch.zeros((len(ds.scene.materials), 5), dtype=torch.float32, device="cuda")
    img = lambda *shape, dt=torch.int32: torch.zeros(shape, dtype=dt, device="cuda")  # noqa: E731
    cam = V.camera_ubo(np.eye(4), np.eye(4), (0.5, 0.5, 0.5))
    out["gbuffer"] = L.vct_raycast_gbuffer(ctx._h, s, C.byref(ds.mesh), surf.data_ptr(), C.byref(cam),
                                           img(H, W).data_ptr(), img(H, W, 4, dt=torch.int16).data_ptr(),
                                           img(H, W).data_ptr(), img(H, W, 4, dt=torch.int16).data_ptr(),
                                           img(H, W, dt=torch.float32).data_ptr(), W, H)
    out["shadow"] = L.vct_raycast_shadow(ctx._h, s, C.byref(ds.mesh), C.byref(ls),
                                         img(H, W, dt=torch.float32).data_ptr(), W, H)
    torch.cuda.synchronize()
    return out


def test_texture_support_and_refusals():
    import torch

    import vk_voxel_cone_tracing_amd as V
    print("[boundary] textured scene through every pass", flush=True)
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    ctx = V.Context(resolution=16, levels=2, radiance_format=V.FORMAT_RGBA8, quirks=V.QUIRK_ALL,
                    downsample_region_size=1, extent_l0=2.0)
    try:
        ctx.update_clip_regions([0.5, 0.5, 0.5])
        plain = _passes(ctx, V.DeviceScene(_scene(False)), s)
        assert all(v == 0 for v in plain.values()), plain
        ds = V.DeviceScene(_scene(True))
        got = _passes(ctx, ds, s)
        assert got == dict(opacity=0, inject=0, svo=0, dense=UNSUPPORTED, gbuffer=UNSUPPORTED, shadow=0), got
        # a texture index past the table
        ds.mesh.texture_count = 3
        got = _passes(ctx, ds, s)
        for k in ("opacity", "inject", "svo"):
            assert got[k] == INVALID, (k, got)
        ds.mesh.texture_count = 4
        # textured materials need texcoords
        ds.mesh.texcoords = None
        got = _passes(ctx, ds, s)
        for k in ("opacity", "inject", "svo"):
            assert got[k] == INVALID, (k, got)
    finally:
        ctx.close()


def test_textured_batch_scene_is_refused():
    import torch

    import vk_voxel_cone_tracing_amd as V
    from vk_voxel_cone_tracing_amd import _abi as A

    print("[boundary] textured batch scene", flush=True)
    ds = V.DeviceScene(_scene(True))
    W = H = 8
    g = dict(diffuse=torch.zeros((H, W), dtype=torch.int32, device="cuda"),
             normal=torch.zeros((H, W, 4), dtype=torch.int16, device="cuda"),
             specular=torch.zeros((H, W), dtype=torch.int32, device="cuda"),
             emission=torch.zeros((H, W, 4), dtype=torch.int16, device="cuda"),
             depth=torch.ones((H, W), dtype=torch.float32, device="cuda"))
    out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    arr = (A.BatchScene * 1)()
    arr[0].mesh = ds.mesh
    arr[0].grid = A.DenseDesc(A.f3(-0.1, -0.1, -0.1), 1.2, 16, 0, 1, 0)
    arr[0].gbuffer = V.gbuffer_struct(g)
    arr[0].camera = A.CameraUBO()
    arr[0].out = out.data_ptr()
    td = A.DenseTraceDesc(0.872665, 1.7, 1.0, 1.0, 2.0, 0.5, 8.0, 0)
    b = C.c_void_p()
    assert A.lib().vct_batch_create(C.byref(b), arr, 1, C.byref(td)) == UNSUPPORTED
    assert not b.value
