"""Python mirror of the reference's per-pass interface over libvct.

RenderPassBase (VFS/RenderPass/RenderPassBase.h:15-73) runs
onBeginRenderPass -> onUpdate -> onEndRenderPass for one FrameLayout; the
RenderPassManager (RenderPassManager.h:19-65) owns the passes, keeps a
name -> resource registry and draws one pass by name. The classes below keep
those names and that order, with a HIP stream in the FrameLayout in place of
the Vulkan command buffer; every GPU operation is one libvct entry point.
"""
from __future__ import annotations

from . import _abi
from . import FORMAT_RGBA16F, Context, Frame, VOLUME_OPACITY, VOLUME_RADIANCE


class RenderPassBase:
    """RenderPassBase::render (RenderPassBase.cpp:60-66)."""

    def __init__(self):
        self.manager: RenderPassManager | None = None

    def attach_render_pass_manager(self, manager: "RenderPassManager"):
        self.manager = manager

    def render(self, frame: Frame):
        self.on_begin_render_pass(frame)
        self.on_update(frame)
        self.on_end_render_pass(frame)

    def on_begin_render_pass(self, frame: Frame):
        pass

    def on_update(self, frame: Frame):
        raise NotImplementedError

    def on_end_render_pass(self, frame: Frame):
        pass


class RenderPassManager:
    """Pass registry plus the string -> resource registry (put/get)."""

    def __init__(self):
        self._passes = {}
        self._resources = {}

    def add_render_pass(self, name: str, render_pass: RenderPassBase):
        render_pass.attach_render_pass_manager(self)
        self._passes[name] = render_pass

    def get_render_pass(self, name: str) -> RenderPassBase:
        return self._passes[name]

    def put(self, name: str, resource):
        self._resources[name] = resource

    def get(self, name: str):
        return self._resources.get(name)

    def draw_single_render_pass(self, name: str, frame: Frame):
        self._passes[name].render(frame)


class VoxelizationPass(RenderPassBase):
    """VoxelizationPass (VoxelizationPass.cpp:74-212): full revoxelization of
    every clip level (the reference forces it, :86), then the opacity mips of
    levels 1..N-1 and the opacity border wrap.

    With `incremental` the pass follows a moving camera the way the
    reference's disabled path does (:89-99, 129-141, 438-494): the first frame
    voxelizes everything; later frames move the regions (`camera_pos` of the
    frame), clear only the slabs that scrolled in (ClipmapCleaner, toroidal)
    and voxelize only those."""

    def __init__(self, ctx: Context, scene, incremental: bool = False):
        super().__init__()
        self.ctx = ctx
        self.scene = scene
        self.incremental = incremental
        self.camera_pos = None
        self._revox = None

    def on_begin_render_pass(self, frame):
        if self.incremental and self._revox is not None and self.camera_pos is not None:
            self._revox = self.ctx.move_clip_regions(self.camera_pos)
            for level, slabs in enumerate(self._revox):
                for mn, ext, _ in slabs:
                    self.ctx.clear_region(VOLUME_OPACITY, mn, ext, level, stream=frame.stream)
            return
        # vkCmdClearColorImage of the opacity atlas (:104-125)
        self.ctx.clear_volume(VOLUME_OPACITY, frame.stream)
        self._revox = [[r] for r in self.ctx.regions]

    def on_update(self, frame):
        if self.incremental:
            for level, slabs in enumerate(self._revox):
                for region in slabs:
                    desc = self.ctx.voxelization_desc(level, region)
                    self.ctx.voxelize_opacity(self.scene, desc, level, frame.stream)
            return
        # every level in one call (vct_voxelize_opacity_levels): the same
        # texels as one vct_voxelize_opacity per level
        descs = [self.ctx.voxelization_desc(level) for level in range(self.ctx.levels)]
        self.ctx.voxelize_opacity_levels(self.scene, descs, frame.stream)

    def on_end_render_pass(self, frame):
        regions = self.ctx.regions
        for level in range(1, self.ctx.levels):
            self.ctx.downsample(VOLUME_OPACITY, level, regions[level - 1][0],
                                self.ctx.config.downsample_region_size, frame.stream)
        self.ctx.wrap_border(VOLUME_OPACITY, frame.stream)


class RadianceInjectionPass(RenderPassBase):
    """RadianceInjectionPass (RadianceInjectionPass.cpp:64-159): level L is
    updated when frameIndex % 2^L == 0 (kUpdateRegionLevelOffsets, :36-38):
    clear, inject, copy alpha from the opacity, then the radiance mips."""

    def __init__(self, ctx: Context, scene, shadow, light_desc, light_shadow, wrap_radiance_border=False):
        super().__init__()
        self.ctx = ctx
        self.scene = scene
        self.shadow = shadow
        self.light_desc = light_desc
        self.light_shadow = light_shadow
        self.frame_index = 0
        # the reference never wraps the radiance border (BorderWrapper.h:33-36 has no caller)
        self.wrap_radiance_border = wrap_radiance_border
        # RGBA16F: vct_begin_radiance_level instead of clear + copy-alpha (same result)
        self.fuse_alpha = True

    def _due(self, level):
        return self.frame_index % (1 << level) == 0

    def _fused(self):
        # half-float radiance: clear + copy-alpha become one pass before injection
        return self.ctx.config.radiance_format == FORMAT_RGBA16F and self.fuse_alpha

    def on_begin_render_pass(self, frame):
        res = self.ctx.resolution
        for level in range(self.ctx.levels):
            if self._due(level):
                if self._fused():
                    self.ctx.begin_radiance_level(level, frame.stream)
                else:
                    self.ctx.clear_region(VOLUME_RADIANCE, (0, 0, 0), (res, res, res), level, 6, frame.stream)

    def on_update(self, frame):
        # the due levels in one call (vct_inject_radiance_levels): one host
        # wait and one sort for all of them, the same texels as one
        # vct_inject_radiance per level
        descs = [self.ctx.voxelization_desc(level) for level in range(self.ctx.levels) if self._due(level)]
        if descs:
            self.ctx.inject_radiance_levels(self.scene, self.shadow, self.light_desc, self.light_shadow, descs,
                                            frame.stream)

    def on_end_render_pass(self, frame):
        for level in range(self.ctx.levels):
            if self._due(level) and not self._fused():
                self.ctx.copy_alpha(level, 6, frame.stream)
        regions = self.ctx.regions
        for level in range(1, self.ctx.levels):
            if self._due(level):
                self.ctx.downsample(VOLUME_RADIANCE, level, regions[level - 1][0],
                                    self.ctx.config.downsample_region_size, frame.stream)
        if self.wrap_radiance_border:
            self.ctx.wrap_border(VOLUME_RADIANCE, frame.stream)
        self.frame_index += 1


class VoxelConeTracingPass(RenderPassBase):
    """VoxelConeTracingPass::onUpdate (VoxelConeTracingPass.cpp:75-106) with
    the member defaults of VoxelConeTracingPass.h:75-82."""

    def __init__(self, ctx: Context, gbuffer, camera, shadow, light_desc, light_shadow, diffuse_out, specular_out,
                 rendering_mode=_abi.MODE_COMBINED_GI):
        super().__init__()
        self.ctx = ctx
        self.gbuffer = gbuffer
        self.camera = camera
        self.shadow = shadow
        self.light_desc = light_desc
        self.light_shadow = light_shadow
        self.diffuse_out = diffuse_out
        self.specular_out = specular_out
        self.rendering_mode = rendering_mode
        self.trace_start_offset = 1.0
        self.indirect_diffuse_intensity = 8.0
        self.ambient_occlusion_factor = 0.5
        self.min_trace_step_factor = 1.0
        self.indirect_specular_intensity = 3.0
        self.occlusion_decay = 2.0
        self.enable_32_cones = False
        self.samples = None  # optional device counter (int64 tensor)

    def desc(self):
        mn, ext, vs = self.ctx.regions[0]
        import numpy as np

        f32 = np.float32
        center = [float(f32(f32(mn[k]) * f32(vs)) + f32(f32(ext[k]) * f32(vs)) * f32(0.5)) for k in range(3)]
        return _abi.ConeTracingDesc(
            _abi.f3(*center), self.rendering_mode, float(vs), float(self.ctx.resolution), self.trace_start_offset,
            self.indirect_diffuse_intensity, self.ambient_occlusion_factor, self.min_trace_step_factor,
            self.indirect_specular_intensity, self.occlusion_decay, int(self.enable_32_cones),
        )

    def on_update(self, frame):
        self.ctx.trace(self.gbuffer, self.camera, self.desc(), self.shadow, self.light_desc, self.light_shadow,
                       self.diffuse_out, self.specular_out, self.samples, frame.stream)
