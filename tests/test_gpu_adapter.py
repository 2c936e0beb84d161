"""The compiled C++ pass adapter (include/vct_passes.hpp, driven by
tests/adapter/clipmap_frame.cpp in a process of its own) against the Python
mirror of the same passes (vk_voxel_cone_tracing_amd/passes.py): one
clipmap GI frame of the textured room (voxelization, injection with mips,
CombinedGI trace) from the same inputs must give identical opacity and
radiance volumes and identical trace images. The Python path is itself
compared with the oracle by the clipmap tests, so this pins the adapter's call sequence and descriptors
(VoxelizationPass.cpp:74-212, RadianceInjectionPass.cpp:64-159,
VoxelConeTracingPass.cpp:75-106)."""
import ctypes as C
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "tests" / "adapter" / "_build" / "clipmap_frame"


def _write_records(path, records):
    with open(path, "wb") as f:
        for name, data in records.items():
            b = bytes(data) if not isinstance(data, np.ndarray) else np.ascontiguousarray(data).tobytes()
            f.write(struct.pack("<I", len(name)) + name.encode() + struct.pack("<Q", len(b)) + b)


def _read_records(path):
    out, data, off = {}, Path(path).read_bytes(), 0
    while off < len(data):
        (n,) = struct.unpack_from("<I", data, off)
        name = data[off + 4: off + 4 + n].decode()
        (size,) = struct.unpack_from("<Q", data, off + 4 + n)
        start = off + 12 + n
        out[name] = data[start: start + size]
        off = start + size
    return out


@pytest.mark.parametrize("fmt", [0, 1])
def test_compiled_adapter_frame_equals_python_passes(tmp_path, fmt):
    import torch

    import vk_voxel_cone_tracing_amd as V
    from vk_voxel_cone_tracing_amd import _abi as A
    from vk_voxel_cone_tracing_amd import passes, scenes, workloads

    assert EXE.exists(), "build the adapter driver first (make -C tests/adapter)"
    print(f"[adapter] fmt={fmt}", flush=True)
    S, W, H, res, levels = 4.0, 96, 64, 32, 3
    scene = scenes.room(2, scale=S, n_boxes=24, n_spheres=4, emissive_boxes=2, textured=True)
    ds = V.DeviceScene(scene)
    # the ray-cast G-buffer takes material factors only: the same room without textures
    plain = V.DeviceScene(scenes.room(2, scale=S, n_boxes=24, n_spheres=4, emissive_boxes=2))
    ctx = V.Context(resolution=res, levels=levels, radiance_format=fmt, quirks=V.QUIRK_ALL, raster_samples=1,
                    downsample_region_size=2, extent_l0=4.0)
    try:
        vp, vp_inv, eye = scenes.camera(2, W, H, scale=S)
        ctx.update_clip_regions(eye)
        light = scenes.light(2, scale=S)
        ld, ls = V.light_structs(light)
        shadow = torch.empty((256, 256), dtype=torch.float32, device="cuda")
        A.call("vct_raycast_shadow", ctx._h, None, C.byref(ds.mesh), C.byref(ls), C.c_void_p(shadow.data_ptr()),
               256, 256)
        g = workloads.gbuffer(ctx._h, plain, plain.scene, (vp, vp_inv, eye), W, H)
        torch.cuda.synchronize()
        sm, gb, cam = V.shadow_struct(shadow), V.gbuffer_struct(g), V.camera_ubo(vp, vp_inv, eye)
        dout = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        sout = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        mgr = passes.RenderPassManager()
        mgr.add_render_pass("VoxelizationPass", passes.VoxelizationPass(ctx, ds))
        mgr.add_render_pass("RadianceInjectionPass", passes.RadianceInjectionPass(ctx, ds, sm, ld, ls))
        mgr.add_render_pass("VoxelConeTracingPass", passes.VoxelConeTracingPass(ctx, gb, cam, sm, ld, ls, dout, sout))
        f = V.Frame()
        for name in ("VoxelizationPass", "RadianceInjectionPass", "VoxelConeTracingPass"):
            mgr.draw_single_render_pass(name, f)
        torch.cuda.synchronize()
        want = dict(opacity=ctx.read_volume(V.VOLUME_OPACITY).tobytes(),
                    radiance=ctx.read_volume(V.VOLUME_RADIANCE).tobytes(),
                    diffuse=dout.cpu().numpy().tobytes(), specular=sout.cpu().numpy().tobytes())
        host = lambda t: t.cpu().numpy()  # noqa: E731
        rec = dict(config=ctx.config, eye=np.asarray(eye, np.float32), positions=scene.positions,
                   normals=scene.normals, texcoords=scene.texcoords, indices=scene.indices, nodes=scene.nodes,
                   materials=scene.materials, primitives=np.ascontiguousarray(scene.primitives, np.uint32), bb_min=scene.bb_min, bb_max=scene.bb_max,
                   shadow=host(shadow), shadow_dims=np.array([256, 256], np.uint32), light=ld, light_shadow=ls,
                   camera=cam, gbuffer_diffuse=host(g["diffuse"]), gbuffer_normal=host(g["normal"]),
                   gbuffer_specular=host(g["specular"]), gbuffer_emission=host(g["emission"]),
                   gbuffer_depth=host(g["depth"]), gbuffer_dims=np.array([W, H], np.uint32))
        rec["texture_count"] = np.array([len(scene.textures)], np.uint32)
        for i, t in enumerate(scene.textures):
            rec[f"texture{i}"] = np.ascontiguousarray(t, np.uint32)
            rec[f"texture{i}_dims"] = np.array([t.shape[1], t.shape[0]], np.uint32)
    finally:
        ctx.close()
    inp, out = tmp_path / "in.rec", tmp_path / "out.rec"
    _write_records(inp, rec)
    r = subprocess.run([str(EXE), str(inp), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout.strip(), r.stderr.strip(), flush=True)
    assert r.returncode == 0, r.stderr
    got = _read_records(out)
    assert got.keys() == want.keys()
    for k in want:
        assert got[k] == want[k], k
    assert np.frombuffer(want["diffuse"], np.float32).any() and np.frombuffer(want["radiance"], np.uint8).any()
