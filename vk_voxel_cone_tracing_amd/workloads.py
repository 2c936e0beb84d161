"""Device set-up of the measured workloads (synthetic, seeded) beside the
1080p clipmap frame: the plain dense VXAO frame, the 512^3 octree build, the
batched dense method against its per-scene loop, and the dense adjoint
against its forward. Every input is produced on the GPU by this package
(procedural scenes, ray-cast G-buffers); timings use HIP events on the
stream the work runs on.
"""
from __future__ import annotations

import ctypes as C
import statistics

import numpy as np
import torch

from . import DeviceScene, _abi, camera_ubo, gbuffer_struct, scenes

A = _abi


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def gbuffer(ctx_h, ds: DeviceScene, scene, cam_triple, W: int, H: int) -> dict:
    """A G-buffer ray cast on the device (vct_raycast_gbuffer)."""
    vp, vp_inv, eye = cam_triple
    g = dict(
        diffuse=torch.empty((H, W), dtype=torch.int32, device="cuda"),
        normal=torch.empty((H, W, 4), dtype=torch.int16, device="cuda"),
        specular=torch.empty((H, W), dtype=torch.int32, device="cuda"),
        emission=torch.empty((H, W, 4), dtype=torch.int16, device="cuda"),
        depth=torch.empty((H, W), dtype=torch.float32, device="cuda"),
    )
    surface = torch.from_numpy(np.ascontiguousarray(scene.surface)).cuda()
    cam = camera_ubo(vp, vp_inv, eye)
    A.call("vct_raycast_gbuffer", ctx_h, _stream(), C.byref(ds.mesh), C.c_void_p(surface.data_ptr()), C.byref(cam),
           C.c_void_p(g["diffuse"].data_ptr()), C.c_void_p(g["normal"].data_ptr()),
           C.c_void_p(g["specular"].data_ptr()), C.c_void_p(g["emission"].data_ptr()),
           C.c_void_p(g["depth"].data_ptr()), W, H)
    torch.cuda.synchronize()
    return g


def scratch_ctx():
    h = C.c_void_p()
    cfg = A.Config(8, 0, 0, A.QUIRK_ALL, 1, 1, 1.0, 0)
    A.call("vct_create", C.byref(h), C.byref(cfg))
    return h


def time_ms(fn, reps: int = 5, warmup: int = 1) -> float:
    """Median of `reps` event-timed calls on the current stream."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        e1.synchronize()
        t.append(e0.elapsed_time(e1))
    return statistics.median(t)


def grid_box(scene):
    lo = np.asarray(scene.bb_min, np.float32)
    ext = float(np.float32((np.asarray(scene.bb_max) - lo).max() * 1.02))
    return [float(v) for v in (lo - np.float32(0.01 * ext))], ext


TRACE_DEFAULTS = dict(aperture=0.872665, min_trace_step_factor=1.0, trace_start_offset=1.0, occlusion_decay=2.0,
                      ambient_occlusion_factor=0.5, indirect_diffuse_intensity=8.0, enable_32_cones=0)


def dense_trace_desc(max_distance: float):
    d = TRACE_DEFAULTS
    return A.DenseTraceDesc(d["aperture"], max_distance, d["min_trace_step_factor"], d["trace_start_offset"],
                            d["occlusion_decay"], d["ambient_occlusion_factor"], d["indirect_diffuse_intensity"],
                            d["enable_32_cones"])


# --------------------------------------------------------------------------
CLIPMAP_FRAME = dict(W=1920, H=1080, scale=16.0, seed=2, res=128, levels=6, shadow=2048)


def clipmap_frame() -> dict:
    """The 1080p clipmap GI frame: the seeded room scaled to 16 units (21 260
    triangles), RGBA16F radiance + RGBA8 opacity clipmaps of 6 x 128^3, a
    2048^2 light depth map ray cast from the seeded light, a 1920x1080
    G-buffer ray cast from the seeded camera with per-pixel perceptual
    roughness U[0.1, 0.8] (metallic > 0 everywhere, so every pixel traces the
    specular cone), and the three passes registered with a RenderPassManager
    (VoxelizationPass, RadianceInjectionPass, VoxelConeTracingPass in
    CombinedGI mode)."""
    from . import Context, FORMAT_RGBA16F, QUIRK_ALL, gbuffer_struct, light_structs, passes, shadow_struct

    F = CLIPMAP_FRAME
    W, H, S = F["W"], F["H"], F["scale"]
    scene = scenes.room(F["seed"], scale=S)
    ds = DeviceScene(scene)
    ctx = Context(resolution=F["res"], levels=F["levels"], radiance_format=FORMAT_RGBA16F, quirks=QUIRK_ALL,
                  raster_samples=1, downsample_region_size=10, extent_l0=16.0)
    vp, vp_inv, eye = scenes.camera(F["seed"], W, H, scale=S)
    ctx.update_clip_regions(eye)
    light = scenes.light(F["seed"], scale=S)
    ld, ls = light_structs(light)
    n = F["shadow"]
    shadow_depth = torch.empty((n, n), dtype=torch.float32, device="cuda")
    A.call("vct_raycast_shadow", ctx._h, _stream(), C.byref(ds.mesh), C.byref(ls),
           C.c_void_p(shadow_depth.data_ptr()), n, n)
    g = gbuffer(ctx._h, ds, scene, (vp, vp_inv, eye), W, H)
    # per-pixel perceptual roughness U[0.1, 0.8] in the diffuse alpha byte
    gen = torch.Generator(device="cuda")
    gen.manual_seed(F["seed"])
    rough = torch.rand((H, W), generator=gen, device="cuda") * 0.7 + 0.1
    rough_b = torch.round(rough * 255.0).to(torch.int32)
    g["diffuse"] = (g["diffuse"] & 0x00FFFFFF) | (rough_b << 24)
    torch.cuda.synchronize()
    sm = shadow_struct(shadow_depth)
    gb = gbuffer_struct(g)
    cam = camera_ubo(vp, vp_inv, eye)
    dout = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    sout = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    mgr = passes.RenderPassManager()
    mgr.add_render_pass("VoxelizationPass", passes.VoxelizationPass(ctx, ds))
    inj = passes.RadianceInjectionPass(ctx, ds, sm, ld, ls)
    mgr.add_render_pass("RadianceInjectionPass", inj)
    vct_pass = passes.VoxelConeTracingPass(ctx, gb, cam, sm, ld, ls, dout, sout)
    vct_pass.samples = torch.zeros(1, dtype=torch.int64, device="cuda")
    mgr.add_render_pass("VoxelConeTracingPass", vct_pass)
    return dict(scene=scene, ds=ds, ctx=ctx, vp=vp, vp_inv=vp_inv, eye=eye, light=light, shadow=shadow_depth, g=g,
                gb=gb, cam=cam, sm=sm, ld=ld, ls=ls, mgr=mgr, inj=inj, vct=vct_pass, W=W, H=H, dout=dout, sout=sout)


def clipmap_frame_run(st: dict, ev=None) -> None:
    """One GI frame through the three passes; with `ev` (4 events) records
    the pass boundaries."""
    from . import Frame

    f = Frame()
    st["inj"].frame_index = 0  # every clip level is due in the measured frame
    st["vct"].samples.zero_()
    for k, name in enumerate(("VoxelizationPass", "RadianceInjectionPass", "VoxelConeTracingPass")):
        if ev is not None:
            ev[k].record()
        st["mgr"].draw_single_render_pass(name, f)
    if ev is not None:
        ev[3].record()


# --------------------------------------------------------------------------
def plain_vxao_inputs(seed: int = 0, res: int = 128, W: int = 640, H: int = 360) -> dict:
    """Inputs of the plain method: the room (about 21k triangles), a ray-cast
    W x H G-buffer from the seeded camera, the res^3 grid box around the
    room and its trace descriptor."""
    scene = scenes.room(seed, scale=1.0)
    ds = DeviceScene(scene)
    vp, vp_inv, eye = scenes.camera(seed, W, H)
    sc = scratch_ctx()
    try:
        g = gbuffer(sc, ds, scene, (vp, vp_inv, eye), W, H)
    finally:
        A.call("vct_destroy", sc)
    gmin, ext = grid_box(scene)
    return dict(scene=scene, ds=ds, g=g, gb=gbuffer_struct(g), vp=vp, vp_inv=vp_inv, eye=eye,
                cam=camera_ubo(vp, vp_inv, eye), gmin=gmin, ext=ext, res=res, W=W, H=H,
                td=dense_trace_desc(float(np.sqrt(3.0)) * ext))


def plain_vxao(seed: int = 0, res: int = 128, W: int = 640, H: int = 360, reps: int = 5) -> dict:
    """The plain method: the room (about 21k triangles) voxelized into a
    res^3 RGBA8 grid, mipped to 1^3, VXAO with 16 cones over a ray-cast W x H
    G-buffer. One frame = clear + voxelize + mips + trace."""
    inp = plain_vxao_inputs(seed, res, W, H)
    scene, ds, gb, cam, gmin, ext, td = (inp[k] for k in ("scene", "ds", "gb", "cam", "gmin", "ext", "td"))
    h = C.c_void_p()
    A.call("vct_dense_create", C.byref(h), C.byref(A.DenseDesc(A.f3(*gmin), ext, res, 0, 1, 0)))
    out = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    samples = torch.zeros(1, dtype=torch.int64, device="cuda")
    s = _stream()

    def frame():
        A.call("vct_dense_clear", h, s)
        A.call("vct_dense_voxelize", h, s, C.byref(ds.mesh))
        A.call("vct_dense_mipmap", h, s)
        A.call("vct_dense_trace", h, s, C.byref(gb), C.byref(cam), C.byref(td), C.c_void_p(out.data_ptr()),
               C.c_void_p(samples.data_ptr()))

    try:
        ms = time_ms(frame, reps)
        samples.zero_()
        frame()
        torch.cuda.synchronize()
        n = int(samples.item())
        trace_ms = time_ms(lambda: A.call("vct_dense_trace", h, s, C.byref(gb), C.byref(cam), C.byref(td),
                                          C.c_void_p(out.data_ptr()), None), reps)
    finally:
        A.call("vct_dense_destroy", h)
    return dict(workload=f"plain VXAO: room seed {seed} ({scene.triangle_count} tris), {res}^3 RGBA8 dense grid "
                         f"mipped to 1^3, 16 cones, {W}x{H}", ms_per_frame=round(ms, 3), trace_ms=round(trace_ms, 3),
                cone_samples=n, cone_samples_per_s=round(n / (trace_ms * 1e-3), 1))


# --------------------------------------------------------------------------
def svo_inputs(seed: int = 1, total_tris: int = 1_000_000, shadow: int = 2048) -> dict:
    """Inputs of the octree build: the room plus thin triangles up to
    `total_tris`, the seeded light and its ray-cast depth map, and a context
    (whose clipmaps are minimal) that owns the fragment list and the pool."""
    from . import Context, light_structs, shadow_struct

    scene = scenes.room(seed, scale=1.0, thin_total=total_tris)
    ds = DeviceScene(scene)
    ctx = Context(resolution=16, levels=1, raster_samples=1, downsample_region_size=1, extent_l0=1.0)
    light = scenes.light(seed, scale=1.0)
    ld, ls = light_structs(light)
    depth = torch.empty((shadow, shadow), dtype=torch.float32, device="cuda")
    A.call("vct_raycast_shadow", ctx._h, _stream(), C.byref(ds.mesh), C.byref(ls), C.c_void_p(depth.data_ptr()),
           shadow, shadow)
    torch.cuda.synchronize()
    return dict(scene=scene, ds=ds, ctx=ctx, light=light, ld=ld, ls=ls, depth=depth, sm=shadow_struct(depth))


def svo_build(seed: int = 1, level: int = 9, total_tris: int = 1_000_000, reps: int = 3) -> dict:
    """The 512^3 sparse voxel octree: the room plus thin triangles up to
    `total_tris`, fragments (K9) then flag / allocate / init per level and the
    leaf and mip writes (K10-K15)."""
    inp = svo_inputs(seed, total_tris)
    scene, ds, ctx, sm, ld, ls = (inp[k] for k in ("scene", "ds", "ctx", "sm", "ld", "ls"))
    s = _stream()
    nfrag = C.c_uint32()

    def build():
        A.call("vct_svo_voxelize", ctx._h, s, C.byref(ds.mesh), C.byref(sm), C.byref(ld), C.byref(ls), level,
               C.byref(nfrag))
        A.call("vct_svo_build", ctx._h, s)

    ms = time_ms(build, reps)
    n = C.c_size_t()
    A.call("vct_svo_read_nodes", ctx._h, s, None, 0, C.byref(n))
    brick_ms = time_ms(lambda: A.call("vct_svo_build_bricks", ctx._h, s), reps)
    return dict(workload=f"SVO {1 << level}^3: room seed {seed} + thin triangles, {scene.triangle_count} tris; "
                         f"node pool and 3^3 RGBA8 bricks",
                ms_per_build=round(ms, 3), ms_bricks=round(brick_ms, 3), ms_build_with_bricks=round(ms + brick_ms, 3),
                fragments=int(nfrag.value), nodes=int(n.value), bricks=int(n.value),
                nodes_per_s=round(n.value / (ms * 1e-3), 1))


# --------------------------------------------------------------------------
def batched_inputs(seed: int = 3, count: int = 64, W: int = 512, H: int = 512):
    """The scenes of the batched workload (scenes.batch) with a ray-cast
    W x H G-buffer each, their grid boxes and output images, and the shared
    trace descriptor. Returns (items, trace descriptor)."""
    items = []
    sc = scratch_ctx()
    try:
        for i, (scene, res) in enumerate(scenes.batch(seed, count)):
            ds = DeviceScene(scene)
            cam3 = scenes.camera(3000 + i, W, H)
            g = gbuffer(sc, ds, scene, cam3, W, H)
            gmin, ext = grid_box(scene)
            items.append(dict(scene=scene, ds=ds, res=res, g=g, gb=gbuffer_struct(g), cam=camera_ubo(*cam3),
                              gmin=gmin, ext=ext, out=torch.empty((H, W, 4), dtype=torch.float32, device="cuda")))
    finally:
        A.call("vct_destroy", sc)
    return items, dense_trace_desc(float(np.sqrt(3.0)))  # the scenes are unit rooms


def batch_scene_array(items):
    """vct_batch_scene entries for batched_inputs' items."""
    arr = (A.BatchScene * len(items))()
    for k, it in enumerate(items):
        arr[k].mesh = it["ds"].mesh
        arr[k].grid = A.DenseDesc(A.f3(*it["gmin"]), it["ext"], it["res"], 0, 1, 0)
        arr[k].gbuffer = it["gb"]
        arr[k].camera = it["cam"]
        arr[k].out = it["out"].data_ptr()
    return arr


def batched(seed: int = 3, count: int = 64, W: int = 512, H: int = 512, reps: int = 3) -> dict:
    """64 scenes of uneven size (5k-500k triangles, 64^3-256^3 grids,
    log-uniform) voxelized, mipped and VXAO-traced at W x H: the per-scene
    loop of vct_dense_* calls against one replay of the batched launch
    (after vct_batch_autotune)."""
    items, td = batched_inputs(seed, count, W, H)
    s = _stream()
    # per-scene loop (grids created once, outside the timing)
    grids = []
    for it in items:
        h = C.c_void_p()
        A.call("vct_dense_create", C.byref(h), C.byref(A.DenseDesc(A.f3(*it["gmin"]), it["ext"], it["res"], 0, 1, 0)))
        grids.append(h)

    def loop(build=True, trace=True):
        for h, it in zip(grids, items):
            if build:
                A.call("vct_dense_clear", h, s)
                A.call("vct_dense_voxelize", h, s, C.byref(it["ds"].mesh))
                A.call("vct_dense_mipmap", h, s)
            if trace:
                A.call("vct_dense_trace", h, s, C.byref(it["gb"]), C.byref(it["cam"]), C.byref(td),
                       C.c_void_p(it["out"].data_ptr()), None)

    try:
        loop_ms = time_ms(loop, reps)
        loop_trace_ms = time_ms(lambda: loop(build=False), reps)  # the loop's 64 traces alone
    finally:
        for h in grids:
            A.call("vct_dense_destroy", h)
    arr = batch_scene_array(items)
    b = C.c_void_p()
    A.call("vct_batch_create", C.byref(b), arr, len(items), C.byref(td))
    try:
        default_ms = time_ms(lambda: A.call("vct_batch_run", b, s, None), reps)
        A.call("vct_batch_set_tuning", b, C.byref(A.BatchTuning(256, 8, 0, 4)))  # the persistent queue
        persistent_ms = time_ms(lambda: A.call("vct_batch_run", b, s, None), reps)
        A.call("vct_batch_status", b)
        best = A.BatchTuning()
        tuned = C.c_float()
        A.call("vct_batch_autotune", b, s, 2, C.byref(best), C.byref(tuned))
        batch_ms = time_ms(lambda: A.call("vct_batch_run", b, s, None), reps)
        A.call("vct_batch_status", b)
    finally:
        A.call("vct_batch_destroy", b)
    tris = sum(it["scene"].triangle_count for it in items)
    return dict(workload=f"{count} scenes, {tris} tris in all, grids 64^3-256^3, VXAO {W}x{H} each",
                per_scene_loop_ms=round(loop_ms, 3), loop_trace_ms=round(loop_trace_ms, 3),
                speedup_bound=round(loop_ms / loop_trace_ms, 3),
                speedup_bound_note="the loop over its own traces: the batched trace runs the same arithmetic",
                batched_default_ms=round(default_ms, 3), batched_persistent_ms=round(persistent_ms, 3),
                batched_ms=round(batch_ms, 3), speedup=round(loop_ms / batch_ms, 3),
                tuning=dict(raster_chunk=best.raster_chunk, mip_brick=best.mip_brick,
                            trace_schedule=best.trace_schedule, workgroups_per_cu=best.workgroups_per_cu))


# --------------------------------------------------------------------------
def adjoint_inputs(seed: int = 4, res: int = 256, views: int = 16, W: int = 512, H: int = 512) -> dict:
    """Inputs of the adjoint workload: `views` ray-cast G-buffers of the room
    (stacked), their camera tensor, a seeded fp32 albedo U[0.2, 0.8] and
    opacity U[0.05, 0.55] grid of res^3 and the target images rendered from
    the grid perturbed by N(0, 0.05) (clamped)."""
    from . import adjoint as adj

    scene = scenes.room(seed, scale=1.0)
    ds = DeviceScene(scene)
    sc = scratch_ctx()
    cams = [scenes.camera(4000 + v, W, H) for v in range(views)]
    try:
        gs = [gbuffer(sc, ds, scene, c, W, H) for c in cams]
    finally:
        A.call("vct_destroy", sc)
    stack = lambda k: torch.stack([g[k] for g in gs]).contiguous()  # noqa: E731
    diffuse, normal, depth = stack("diffuse"), stack("normal"), stack("depth")
    cam_t = adj.cameras_tensor(cams)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    albedo = torch.rand((res, res, res, 3), generator=gen, device="cuda") * 0.6 + 0.2
    opacity = torch.rand((res, res, res), generator=gen, device="cuda") * 0.5 + 0.05
    p = adj.TraceParams(max_distance=float(np.sqrt(3.0)))
    with torch.no_grad():
        ta = (albedo + 0.05 * torch.randn(albedo.shape, generator=gen, device="cuda")).clamp(0.0, 1.0)
        to = (opacity + 0.05 * torch.randn(opacity.shape, generator=gen, device="cuda")).clamp(0.0, 0.95)
        target = adj.render(ta, to, diffuse, normal, depth, cam_t, (0.0, 0.0, 0.0), 1.0, p)
    return dict(scene=scene, ds=ds, gbuffers=gs, cams=cams, diffuse=diffuse, normal=normal, depth=depth,
                cameras=cam_t, albedo=albedo, opacity=opacity, target=target, params=p, res=res, W=W, H=H)


def adjoint(seed: int = 4, res: int = 256, views: int = 16, W: int = 512, H: int = 512, reps: int = 3) -> dict:
    """fp32 res^3 albedo + opacity grid, `views` seeded views: forward render
    and the L2 loss against targets rendered from a perturbed grid
    (sigma 0.05), then the backward through cone marching and mips."""
    from . import adjoint as adj

    inp = adjoint_inputs(seed, res, views, W, H)
    diffuse, normal, depth, cam_t, albedo, opacity, target, p = (
        inp[k] for k in ("diffuse", "normal", "depth", "cameras", "albedo", "opacity", "target", "params"))
    a = albedo.clone().requires_grad_(True)
    o = opacity.clone().requires_grad_(True)

    def fwd():
        with torch.no_grad():
            out = adj.render(a, o, diffuse, normal, depth, cam_t, (0.0, 0.0, 0.0), 1.0, p)
            return ((out - target) ** 2).sum()

    res_out = {}
    fwd_ms = time_ms(fwd, reps)

    def fwd_tape():  # the forward as a training step runs it: grad mode, so it writes the tape
        out = adj.render(a, o, diffuse, normal, depth, cam_t, (0.0, 0.0, 0.0), 1.0, p)
        return ((out - target) ** 2).sum()

    fwd_tape_ms = time_ms(fwd_tape, reps)
    for mode in ("tape", "keep", "recompute"):
        def step():
            a.grad = None
            o.grad = None
            out = adj.render(a, o, diffuse, normal, depth, cam_t, (0.0, 0.0, 0.0), 1.0, p,
                             recompute=mode == "recompute", tape=mode == "tape")
            ((out - target) ** 2).sum().backward()

        res_out[mode] = round(time_ms(step, reps), 3)
    best = min(res_out.values())
    return dict(workload=f"adjoint: fp32 {res}^3 albedo+opacity, {views} views {W}x{H}, L2 vs targets from a "
                         f"perturbed grid (sigma 0.05)", forward_ms=round(fwd_ms, 3),
                forward_with_tape_ms=round(fwd_tape_ms, 3), forward_backward_ms=res_out,
                backward_over_forward=round((best - fwd_ms) / fwd_ms, 3),
                backward_over_forward_note="(forward + backward in grad mode, best of the three sample modes, "
                                           "minus the no-grad forward) / the no-grad forward")
