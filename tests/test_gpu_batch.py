"""The batched dense method (one flat launch per phase, or one persistent
launch; either replayed from a hipGraph) against the per-scene dense calls it
batches.

Every scene's mip pyramid must be bit-identical to vct_dense_voxelize +
vct_dense_mipmap on the same scene, the traced images must equal
vct_dense_trace's (the kernels run the same arithmetic), the executed cone
samples must add up to the per-scene total, and all of that must hold on a
second replay (winners and counters are reset by the launch itself) and
under every schedule the autotuner can pick.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_api as O

pytestmark = pytest.mark.gpu

W, H = 56, 40  # not a multiple of the 16-pixel tile


def _scenes():
    from vk_voxel_cone_tracing_amd import scenes

    specs = [(16, 2, 0, 0), (32, 8, 1, 3000), (64, 24, 4, 9000), (32, 4, 0, 0), (16, 40, 2, 12000), (64, 10, 1, 0)]
    out = []
    for i, (res, boxes, spheres, thin) in enumerate(specs):
        sc = scenes.room(500 + i, scale=1.0, n_boxes=boxes, n_spheres=spheres, thin_total=thin)
        out.append((sc, res))
    return out


def _setup():
    import torch
    import vk_voxel_cone_tracing_amd as V
    from vk_voxel_cone_tracing_amd import scenes

    items = []
    for i, (sc, res) in enumerate(_scenes()):
        lo = np.asarray(sc.bb_min, np.float32)
        ext = float(np.float32((np.asarray(sc.bb_max) - lo).max() * 1.02))
        gmin = [float(v) for v in (lo - np.float32(0.01 * ext))]
        vp, vp_inv, eye = scenes.camera(600 + i, W, H, scale=1.0)
        pos, nrm, mat = sc.world_triangles()
        g = O.raycast_gbuffer(sc, pos, nrm, mat, vp, vp_inv, W, H)
        gd = {k: torch.from_numpy(v.view(np.int32) if v.dtype == np.uint32 else
                                  (v.view(np.int16) if v.dtype == np.uint16 else v)).cuda() for k, v in g.items()}
        items.append(dict(scene=sc, ds=V.DeviceScene(sc), res=res, gmin=gmin, ext=ext,
                          cam=V.camera_ubo(vp, vp_inv, eye), gd=gd, gb=V.gbuffer_struct(gd)))
    return items


TD = dict(aperture=0.872665, max_distance=1.2, min_trace_step_factor=1.0, trace_start_offset=1.0,
          occlusion_decay=2.0, ambient_occlusion_factor=0.5, indirect_diffuse_intensity=8.0, enable_32_cones=0)


def _td():
    from vk_voxel_cone_tracing_amd import _abi

    return _abi.DenseTraceDesc(*TD.values())


def _read_pyramid(read, res):
    levels = []
    for m in range(int(np.log2(res)) + 1):
        n = (res >> m) + 2
        a = np.zeros((1 if m == 0 else 6, n, n, n), np.uint32)
        read(m, a)
        levels.append(a)
    return levels


def _plain(items):
    import torch
    import vk_voxel_cone_tracing_amd as V

    A = V._abi
    outs, pyrs, total = [], [], 0
    for it in items:
        desc = A.DenseDesc(A.f3(*it["gmin"]), it["ext"], it["res"], 0, 1, 0)
        h = C.c_void_p()
        A.call("vct_dense_create", C.byref(h), C.byref(desc))
        try:
            out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
            smp = torch.zeros(1, dtype=torch.int64, device="cuda")
            A.call("vct_dense_clear", h, None)
            A.call("vct_dense_voxelize", h, None, C.byref(it["ds"].mesh))
            A.call("vct_dense_mipmap", h, None)
            A.call("vct_dense_trace", h, None, C.byref(it["gb"]), C.byref(it["cam"]), C.byref(_td()),
                   C.c_void_p(out.data_ptr()), C.c_void_p(smp.data_ptr()))
            torch.cuda.synchronize()
            pyrs.append(_read_pyramid(
                lambda m, a: A.call("vct_dense_read_level", h, None, m, a.ctypes.data_as(C.c_void_p), a.nbytes),
                it["res"]))
            outs.append(out.cpu().numpy())
            total += int(smp.item())
        finally:
            A.call("vct_dense_destroy", h)
    return outs, pyrs, total


def _make_batch(items, outs):
    import vk_voxel_cone_tracing_amd as V

    A = V._abi
    arr = (A.BatchScene * len(items))()
    for k, it in enumerate(items):
        arr[k].mesh = it["ds"].mesh
        arr[k].grid = A.DenseDesc(A.f3(*it["gmin"]), it["ext"], it["res"], 0, 1, 0)
        arr[k].gbuffer = it["gb"]
        arr[k].camera = it["cam"]
        arr[k].out = outs[k].data_ptr()
    h = C.c_void_p()
    A.call("vct_batch_create", C.byref(h), arr, len(items), C.byref(_td()))
    return h


def _run_batch(h, items, outs):
    import torch
    import vk_voxel_cone_tracing_amd as V

    A = V._abi
    for o in outs:
        o.fill_(-1.0)
    smp = torch.zeros(1, dtype=torch.int64, device="cuda")
    A.call("vct_batch_run", h, None, C.c_void_p(smp.data_ptr()))
    torch.cuda.synchronize()
    A.call("vct_batch_status", h)
    pyrs = [_read_pyramid(lambda m, a, k=k: A.call("vct_batch_read_level", h, None, k, m,
                                                   a.ctypes.data_as(C.c_void_p), a.nbytes), it["res"])
            for k, it in enumerate(items)]
    return [o.cpu().numpy() for o in outs], pyrs, int(smp.item())


def test_batch_matches_per_scene_calls():
    import torch
    import vk_voxel_cone_tracing_amd as V

    A = V._abi
    items = _setup()
    ref_out, ref_pyr, ref_total = _plain(items)
    outs = [torch.zeros((H, W, 4), dtype=torch.float32, device="cuda") for _ in items]
    h = _make_batch(items, outs)
    try:
        # workgroups_per_cu 0: one flat launch per phase; 1..4: the persistent queue
        schedules = [(256, 8, 0, 0), (256, 8, 0, 0), (64, 4, 1, 0), (256, 8, 0, 2), (64, 4, 1, 1), (128, 8, 1, 4)]
        for rc, mb, sched, occ in schedules:
            A.call("vct_batch_set_tuning", h, C.byref(A.BatchTuning(rc, mb, sched, occ)))
            got_out, got_pyr, total = _run_batch(h, items, outs)
            print(f"tuning ({rc},{mb},{sched},{occ}): samples {total} (per-scene {ref_total})")
            for k in range(len(items)):
                occ_count = int((ref_pyr[k][0] != 0).sum())
                assert occ_count > 50, (k, occ_count)
                for m, (a, b) in enumerate(zip(got_pyr[k], ref_pyr[k])):
                    np.testing.assert_array_equal(a, b, err_msg=f"scene {k} level {m}")
                err = np.abs(got_out[k] - ref_out[k]).max()
                assert err <= 1e-6 * max(1.0, np.abs(ref_out[k]).max()), (k, err)
            assert total == ref_total
        best = A.BatchTuning()
        ms = C.c_float()
        A.call("vct_batch_autotune", h, None, 1, C.byref(best), C.byref(ms))
        print(f"autotuned ({best.raster_chunk},{best.mip_brick},{best.trace_schedule},{best.workgroups_per_cu}) "
              f"{ms.value:.3f} ms")
        assert best.raster_chunk in (64, 128, 256) and best.mip_brick in (4, 8)
        got_out, got_pyr, total = _run_batch(h, items, outs)
        for k in range(len(items)):
            np.testing.assert_array_equal(got_pyr[k][-1], ref_pyr[k][-1])
            assert np.abs(got_out[k] - ref_out[k]).max() <= 1e-6 * max(1.0, np.abs(ref_out[k]).max())
        assert total == ref_total
    finally:
        A.call("vct_batch_destroy", h)


def test_batch_rejects_bad_scenes_without_a_launch():
    import torch
    import vk_voxel_cone_tracing_amd as V

    A = V._abi
    items = _setup()[:1]
    outs = [torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")]
    arr = (A.BatchScene * 1)()
    arr[0].mesh = items[0]["ds"].mesh
    arr[0].grid = A.DenseDesc(A.f3(*items[0]["gmin"]), items[0]["ext"], 48, 0, 1, 0)  # not a power of two
    arr[0].gbuffer = items[0]["gb"]
    arr[0].camera = items[0]["cam"]
    arr[0].out = outs[0].data_ptr()
    h = C.c_void_p()
    assert A.lib().vct_batch_create(C.byref(h), arr, 1, C.byref(_td())) == A.VCT_ERR_INVALID_ARGUMENT
    arr[0].grid = A.DenseDesc(A.f3(*items[0]["gmin"]), items[0]["ext"], 32, 2, 1, 0)  # RGBA32F
    assert A.lib().vct_batch_create(C.byref(h), arr, 1, C.byref(_td())) == A.VCT_ERR_UNSUPPORTED
