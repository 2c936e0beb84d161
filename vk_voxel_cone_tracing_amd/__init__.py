"""vk_voxel_cone_tracing_amd - MI355X (gfx950) voxel cone tracing hot path.

The product is libvct.so (HIP kernels behind the C ABI in include/vct.h).
This package loads it with ctypes and offers:

* `Context`: one vct_ctx (clipmap volumes + scratch), one method per C entry
  point, device buffers passed as torch CUDA tensors;
* `passes`: Python mirrors of the reference's render passes
  (VoxelizationPass, RadianceInjectionPass, VoxelConeTracingPass) driving a
  Context frame by frame, as RenderPassManager drives the Vulkan passes;
* `scenes` / `glm`: seeded procedural inputs and glm-convention matrices.

Nothing here computes on the CPU in place of the library: if libvct.so is
missing every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import itertools
from dataclasses import dataclass

import numpy as np

from . import _abi
from ._abi import (  # noqa: F401
    FORMAT_RGBA8,
    FORMAT_RGBA16F,
    MODE_AMBIENT_OCCLUSION,
    MODE_COMBINED_GI,
    QUIRK_ALL,
    VOLUME_OPACITY,
    VOLUME_RADIANCE,
    VctError,
)

__all__ = ["Context", "DeviceScene", "Frame", "lib", "version"]


def lib():
    return _abi.lib()


def version() -> str:
    return _abi.lib().vct_version().decode()


def _stream_handle(stream) -> C.c_void_p:
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _ptr(t) -> C.c_void_p:
    return C.c_void_p(0 if t is None else t.data_ptr())


@dataclass
class Frame:
    """FrameLayout (VulkanFramework/FrameLayout.h:10-16) with a HIP stream in
    place of the command buffer."""

    stream: object = None
    frame_index: int = 0
    frame_time: float = 0.0


_MATERIAL_VERSIONS = itertools.count(1)


class DeviceScene:
    """A Scene uploaded to device memory plus the vct_mesh describing it."""

    def __init__(self, scene, device="cuda"):
        import torch

        self.scene = scene
        self.positions = torch.from_numpy(np.ascontiguousarray(scene.positions)).to(device)
        self.normals = torch.from_numpy(np.ascontiguousarray(scene.normals)).to(device)
        self.texcoords = torch.from_numpy(np.ascontiguousarray(scene.texcoords)).to(device)
        self.indices = torch.from_numpy(np.ascontiguousarray(scene.indices).view(np.int32)).to(device)
        self.nodes = torch.from_numpy(np.ascontiguousarray(scene.nodes)).to(device)
        self.materials = torch.from_numpy(np.ascontiguousarray(scene.materials).view(np.uint8)).to(device)
        prims = scene.primitives
        self._prims = (_abi.Primitive * len(prims))()
        for i, (first, count, voff, inst, mat) in enumerate(prims):
            self._prims[i] = _abi.Primitive(int(first), int(count), int(np.int32(np.uint32(voff))), int(inst), int(mat))
        m = _abi.Mesh()
        m.positions = self.positions.data_ptr()
        m.normals = self.normals.data_ptr()
        m.texcoords = self.texcoords.data_ptr()
        m.indices = self.indices.data_ptr()
        m.node_matrices = self.nodes.data_ptr()
        m.materials = self.materials.data_ptr()
        m.vertex_count = len(scene.positions)
        m.index_count = len(scene.indices)
        m.instance_count = len(scene.nodes)
        m.material_count = len(scene.materials)
        m.primitives = C.cast(self._prims, C.POINTER(_abi.Primitive))
        m.primitive_count = len(prims)
        m.scene_bb_min = _abi.f3(*scene.bb_min.tolist())
        m.scene_bb_max = _abi.f3(*scene.bb_max.tolist())
        textures = getattr(scene, "textures", [])
        self.textures = [torch.from_numpy(np.ascontiguousarray(t, np.uint32).view(np.int32)).to(device)
                         for t in textures]
        self._tex = (_abi.Texture * max(1, len(textures)))()
        for i, (t, d) in enumerate(zip(textures, self.textures)):
            self._tex[i] = _abi.Texture(d.data_ptr(), int(t.shape[1]), int(t.shape[0]))
        m.textures = C.cast(self._tex, C.POINTER(_abi.Texture)) if textures else None
        m.texture_count = len(textures)
        # a new material table each time: its address may be a freed one's
        m.material_version = next(_MATERIAL_VERSIONS)
        self.mesh = m


def light_structs(light: dict):
    from . import glm

    ld = _abi.DirectionalLightDesc()
    ld.direction = _abi.f3(*np.asarray(light["direction"], np.float32).tolist())
    ld.intensity = float(light["intensity"])
    ld.color = _abi.f3(*np.asarray(light["color"], np.float32).tolist())
    ls = _abi.DirectionalLightShadowDesc()
    ls.view = _abi.m4(*glm.flat(light["view"]).tolist())
    ls.proj = _abi.m4(*glm.flat(light["proj"]).tolist())
    ls.z_near = float(light["z_near"])
    ls.z_far = float(light["z_far"])
    return ld, ls


def camera_ubo(view_proj, view_proj_inv, eye):
    from . import glm

    cam = _abi.CameraUBO()
    cam.view_proj = _abi.m4(*glm.flat(view_proj).tolist())
    cam.view_proj_inv = _abi.m4(*glm.flat(view_proj_inv).tolist())
    cam.eye_pos = _abi.f3(*np.asarray(eye, np.float32).tolist())
    return cam


def shadow_struct(depth_tensor):
    """vct_shadow_map over a device tensor; the struct keeps the tensor alive,
    so the memory it points to cannot be recycled while the struct exists."""
    h, w = depth_tensor.shape
    s = _abi.ShadowMap(depth_tensor.data_ptr(), w, h)
    s._keep = depth_tensor
    return s


def gbuffer_struct(g: dict):
    """vct_gbuffer over device tensors (kept alive by the struct)."""
    h, w = g["depth"].shape
    s = _abi.GBuffer(
        g["diffuse"].data_ptr(), g["normal"].data_ptr(), g["specular"].data_ptr(), g["emission"].data_ptr(),
        g["depth"].data_ptr(), w, h,
    )
    s._keep = dict(g)
    return s


class Context:
    """One vct_ctx: the opacity and radiance clipmaps of `levels` x res^3 texels
    (6 faces, 1-texel border) and the scratch of every pass."""

    def __init__(self, resolution=128, levels=6, radiance_format=FORMAT_RGBA8, quirks=QUIRK_ALL,
                 raster_samples=1, downsample_region_size=10, extent_l0=16.0):
        L = _abi.lib()
        cfg = _abi.Config(resolution, levels, radiance_format, quirks, raster_samples, downsample_region_size,
                          float(extent_l0), 0)
        self.config = cfg
        self._h = C.c_void_p()
        _abi.check("vct_create", L.vct_create(C.byref(self._h), C.byref(cfg)))
        self.resolution = resolution
        self.levels = levels
        self.radiance_format = radiance_format
        self.quirks = quirks

    def close(self):
        if self._h:
            _abi.lib().vct_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- regions
    @property
    def regions(self):
        arr = (_abi.ClipmapRegion * self.levels)()
        _abi.call("vct_get_clip_regions", self._h, arr, self.levels)
        return [(tuple(r.min_corner), tuple(r.extent), r.voxel_size) for r in arr]

    def set_regions(self, regions):
        arr = (_abi.ClipmapRegion * self.levels)()
        for i, (mn, ext, vs) in enumerate(regions):
            arr[i] = _abi.ClipmapRegion(_abi.i3(*mn), _abi.u3(*ext), float(vs))
        _abi.call("vct_set_clip_regions", self._h, arr, self.levels)

    def update_clip_regions(self, camera_pos):
        cp = (C.c_float * 3)(*[float(x) for x in camera_pos])
        _abi.call("vct_update_clip_regions", self._h, cp)

    def move_clip_regions(self, camera_pos):
        """vct_move_clip_regions: moves the regions with the camera and returns,
        per level, the regions to clear and voxelize again."""
        cp = (C.c_float * 3)(*[float(x) for x in camera_pos])
        n = _abi.REVOX_PER_LEVEL
        revox = (_abi.ClipmapRegion * (n * self.levels))()
        counts = (C.c_uint32 * self.levels)()
        _abi.call("vct_move_clip_regions", self._h, cp, revox, counts)
        return [[(tuple(r.min_corner), tuple(r.extent), r.voxel_size) for r in revox[n * i: n * i + counts[i]]]
                for i in range(self.levels)]

    def voxelization_desc(self, level, region=None):
        d = _abi.VoxelizationDesc()
        rp = None
        if region is not None:
            mn, ext, vs = region
            rp = C.byref(_abi.ClipmapRegion(_abi.i3(*mn), _abi.u3(*ext), float(vs)))
        _abi.call("vct_make_voxelization_desc", self._h, level, rp, C.byref(d))
        return d

    # ---- passes
    def clear_volume(self, volume, stream=None):
        _abi.call("vct_clear_volume", self._h, _stream_handle(stream), volume)

    def voxelize_opacity(self, dscene: DeviceScene, desc, level, stream=None):
        _abi.call("vct_voxelize_opacity", self._h, _stream_handle(stream), C.byref(dscene.mesh), C.byref(desc), level)

    def inject_radiance(self, dscene: DeviceScene, shadow, light_desc, light_shadow, desc, level, stream=None):
        _abi.call("vct_inject_radiance", self._h, _stream_handle(stream), C.byref(dscene.mesh), C.byref(shadow),
                  C.byref(light_desc), C.byref(light_shadow), C.byref(desc), level)

    def voxelize_opacity_levels(self, dscene: DeviceScene, descs, stream=None):
        """vct_voxelize_opacity_levels: the descriptors' clip levels in one call."""
        arr = (_abi.VoxelizationDesc * len(descs))(*descs)
        _abi.call("vct_voxelize_opacity_levels", self._h, _stream_handle(stream), C.byref(dscene.mesh), arr, len(descs))

    def inject_radiance_levels(self, dscene: DeviceScene, shadow, light_desc, light_shadow, descs, stream=None):
        """vct_inject_radiance_levels: the descriptors' clip levels in one call."""
        arr = (_abi.VoxelizationDesc * len(descs))(*descs)
        _abi.call("vct_inject_radiance_levels", self._h, _stream_handle(stream), C.byref(dscene.mesh),
                  C.byref(shadow), C.byref(light_desc), C.byref(light_shadow), arr, len(descs))

    def clear_region(self, volume, min_corner, extent, level, face_count=6, stream=None):
        d = _abi.ImageCleaningDesc(_abi.i3(*min_corner), level, _abi.u3(*extent), self.resolution, face_count)
        _abi.call("vct_clear_region", self._h, _stream_handle(stream), volume, C.byref(d))

    def downsample(self, volume, level, prev_min_corner, region_size=10, stream=None):
        d = _abi.DownSampleDesc(_abi.i3(*prev_min_corner), level, self.resolution, region_size)
        _abi.call("vct_downsample", self._h, _stream_handle(stream), volume, C.byref(d))

    def begin_radiance_level(self, level, stream=None):
        _abi.call("vct_begin_radiance_level", self._h, _stream_handle(stream), level)

    def copy_alpha(self, level, face_count=6, stream=None):
        d = _abi.CopyAlphaDesc(level, self.resolution, face_count)
        _abi.call("vct_copy_alpha", self._h, _stream_handle(stream), C.byref(d))

    def wrap_border(self, volume, stream=None):
        d = _abi.BorderWrappingDesc(self.resolution, 2, 6, self.levels)
        _abi.call("vct_wrap_border", self._h, _stream_handle(stream), volume, C.byref(d))

    def trace(self, gbuffer, camera, desc, shadow, light_desc, light_shadow, diffuse_out, specular_out,
              samples=None, stream=None):
        _abi.call(
            "vct_trace", self._h, _stream_handle(stream), C.byref(gbuffer), C.byref(camera), C.byref(desc),
            C.byref(shadow) if shadow is not None else None,
            C.byref(light_desc) if light_desc is not None else None,
            C.byref(light_shadow) if light_shadow is not None else None,
            _ptr(diffuse_out), _ptr(specular_out), _ptr(samples),
        )

    # ---- volume access (reference atlas layout)
    def volume_shape(self, volume):
        D = self.resolution + 2
        comps = (4, np.uint16) if (volume == VOLUME_RADIANCE and self.radiance_format == FORMAT_RGBA16F) else (1, np.uint32)
        return (D, D * self.levels, D * 6, comps[0]), comps[1]

    def read_volume(self, volume, stream=None) -> np.ndarray:
        shape, dt = self.volume_shape(volume)
        out = np.empty(shape, dt)
        _abi.call("vct_read_volume", self._h, _stream_handle(stream), volume, out.ctypes.data_as(C.c_void_p), out.nbytes)
        return out if shape[3] > 1 else out[..., 0]

    def write_volume(self, volume, arr, stream=None):
        shape, dt = self.volume_shape(volume)
        a = np.ascontiguousarray(arr, dtype=dt).reshape(shape)
        _abi.call("vct_write_volume", self._h, _stream_handle(stream), volume, a.ctypes.data_as(C.c_void_p), a.nbytes)
