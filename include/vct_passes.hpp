// vct_passes.hpp - the reference's clipmap render passes as C++ classes over
// the C ABI of vct.h (header only, C++17).
//
// The reference records each pass into a Vulkan command buffer between
// RenderPassBase::beginRenderPass and endRenderPass (RenderPassBase.h:15-73,
// RenderPassBase.cpp:60-66: render() = onBeginRenderPass, onUpdate,
// onEndRenderPass). Here a FrameLayout carries a HIP stream instead of the
// command buffer and every recorded dispatch becomes one vct_* call on that
// stream; stream order replaces the pipeline barriers. A renderer keeps its
// pass classes and registration and swaps their bodies for these.
//
//   VoxelizationPassHip      VoxelizationPass.cpp:74-212
//   RadianceInjectionPassHip RadianceInjectionPass.cpp:64-159
//   VoxelConeTracingPassHip  VoxelConeTracingPass.cpp:75-106
//   RenderPassManagerHip     RenderPassManager.h:19-65 (passes by name)
//
// Errors: every failing call throws vct::PassError with the library's
// vct_last_error() text.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>

#include "vct.h"

namespace vct {

struct PassError : std::runtime_error {
    vct_status status;
    PassError(vct_status s, const std::string& what) : std::runtime_error(what), status(s) {}
};

inline void check(vct_status s, const char* call)
{
    if (s != VCT_OK) throw PassError(s, std::string(call) + ": " + vct_status_string(s) + ": " + vct_last_error());
}

#define VCT_PASS_CHECK(expr) ::vct::check((expr), #expr)

// FrameLayout (FrameLayout.h:10-16) with the stream the frame records on.
struct FrameLayout {
    hipStream_t stream = nullptr;
    uint32_t frame_index = 0;
    float frame_time = 0.0f;
};

// RenderPassBase (RenderPassBase.h:15-73).
class RenderPassBase {
public:
    virtual ~RenderPassBase() = default;
    void render(const FrameLayout* frame)
    {
        onBeginRenderPass(frame);
        onUpdate(frame);
        onEndRenderPass(frame);
    }

protected:
    virtual void onBeginRenderPass(const FrameLayout*) {}
    virtual void onEndRenderPass(const FrameLayout*) {}
    virtual void onUpdate(const FrameLayout*) = 0;
};

// RenderPassManager (RenderPassManager.h:19-65): owns the passes and draws
// one by name.
class RenderPassManagerHip {
public:
    template <typename Pass, typename... Args>
    Pass* addRenderPass(const std::string& name, Args&&... args)
    {
        auto p = std::make_unique<Pass>(std::forward<Args>(args)...);
        Pass* raw = p.get();
        _passes[name] = std::move(p);
        return raw;
    }
    RenderPassBase* getRenderPass(const std::string& name) const
    {
        auto it = _passes.find(name);
        return it == _passes.end() ? nullptr : it->second.get();
    }
    void drawSingleRenderPass(const std::string& name, const FrameLayout* frame)
    {
        RenderPassBase* p = getRenderPass(name);
        if (!p) throw std::out_of_range("no render pass named " + name);
        p->render(frame);
    }

private:
    std::map<std::string, std::unique_ptr<RenderPassBase>> _passes;
};

// The context's configuration and current clip regions.
inline vct_clipmap_region clip_region(const vct_ctx* ctx, uint32_t level, uint32_t levels)
{
    vct_clipmap_region r[8];
    VCT_PASS_CHECK(vct_get_clip_regions(ctx, r, levels));
    return r[level];
}

// VoxelizationPass (VoxelizationPass.cpp:74-212) with the reference's forced
// full revoxelization (:86): clear the opacity atlas, voxelize every clip
// level, then the opacity mips of levels 1..N-1 (DownSampler.cpp:115-168) and
// the border wrap (BorderWrapper.cpp:69-73).
class VoxelizationPassHip : public RenderPassBase {
public:
    VoxelizationPassHip(vct_ctx* ctx, const vct_config& cfg, const vct_mesh* mesh)
        : _ctx(ctx), _cfg(cfg), _mesh(mesh)
    {
    }

protected:
    void onBeginRenderPass(const FrameLayout* f) override
    {
        VCT_PASS_CHECK(vct_clear_volume(_ctx, f->stream, VCT_VOLUME_OPACITY));
    }
    void onUpdate(const FrameLayout* f) override
    {
        vct_voxelization_desc descs[8];
        for (uint32_t level = 0; level < _cfg.clip_levels; ++level) {
            const vct_clipmap_region r = clip_region(_ctx, level, _cfg.clip_levels);
            VCT_PASS_CHECK(vct_make_voxelization_desc(_ctx, level, &r, &descs[level]));
        }
        // every level in one call: the texels of one vct_voxelize_opacity per level
        VCT_PASS_CHECK(vct_voxelize_opacity_levels(_ctx, f->stream, _mesh, descs, _cfg.clip_levels));
    }
    void onEndRenderPass(const FrameLayout* f) override
    {
        for (uint32_t level = 1; level < _cfg.clip_levels; ++level) {
            const vct_clipmap_region prev = clip_region(_ctx, level - 1, _cfg.clip_levels);
            vct_downsample_desc d{{prev.min_corner[0], prev.min_corner[1], prev.min_corner[2]},
                                  (int32_t)level,
                                  (int32_t)_cfg.voxel_resolution,
                                  (int32_t)_cfg.downsample_region_size};
            VCT_PASS_CHECK(vct_downsample(_ctx, f->stream, VCT_VOLUME_OPACITY, &d));
        }
        vct_border_wrapping_desc bw{_cfg.voxel_resolution, 2u, 6u, _cfg.clip_levels};
        VCT_PASS_CHECK(vct_wrap_border(_ctx, f->stream, VCT_VOLUME_OPACITY, &bw));
    }

private:
    vct_ctx* _ctx;
    vct_config _cfg;
    const vct_mesh* _mesh;
};

// RadianceInjectionPass (RadianceInjectionPass.cpp:64-159): clip level L is
// refreshed when frame_index % 2^L == 0 (:36-38): clear, inject, copy the
// opacity alpha, then the radiance mips of the refreshed levels. With RGBA16F
// radiance the clear and the copy-alpha are one vct_begin_radiance_level
// before the injection (bit-identical result).
class RadianceInjectionPassHip : public RenderPassBase {
public:
    RadianceInjectionPassHip(vct_ctx* ctx, const vct_config& cfg, const vct_mesh* mesh, const vct_shadow_map* shadow,
                             const vct_directional_light_desc* light,
                             const vct_directional_light_shadow_desc* light_shadow)
        : _ctx(ctx), _cfg(cfg), _mesh(mesh), _shadow(shadow), _light(light), _light_shadow(light_shadow)
    {
    }

protected:
    bool due(const FrameLayout* f, uint32_t level) const { return f->frame_index % (1u << level) == 0; }
    bool fused() const { return _cfg.radiance_format == VCT_FORMAT_RGBA16F; }

    void onBeginRenderPass(const FrameLayout* f) override
    {
        const int32_t res = (int32_t)_cfg.voxel_resolution;
        for (uint32_t level = 0; level < _cfg.clip_levels; ++level) {
            if (!due(f, level)) continue;
            if (fused()) {
                VCT_PASS_CHECK(vct_begin_radiance_level(_ctx, f->stream, level));
            } else {
                vct_image_cleaning_desc d{{0, 0, 0}, (int32_t)level, {(uint32_t)res, (uint32_t)res, (uint32_t)res},
                                          res, 6u};
                VCT_PASS_CHECK(vct_clear_region(_ctx, f->stream, VCT_VOLUME_RADIANCE, &d));
            }
        }
    }
    void onUpdate(const FrameLayout* f) override
    {
        vct_voxelization_desc descs[8];
        uint32_t n = 0;
        for (uint32_t level = 0; level < _cfg.clip_levels; ++level) {
            if (!due(f, level)) continue;
            const vct_clipmap_region r = clip_region(_ctx, level, _cfg.clip_levels);
            VCT_PASS_CHECK(vct_make_voxelization_desc(_ctx, level, &r, &descs[n++]));
        }
        // the due levels in one call: one host wait and one sort for all of them
        if (n)
            VCT_PASS_CHECK(
                vct_inject_radiance_levels(_ctx, f->stream, _mesh, _shadow, _light, _light_shadow, descs, n));
    }
    void onEndRenderPass(const FrameLayout* f) override
    {
        for (uint32_t level = 0; level < _cfg.clip_levels && !fused(); ++level) {
            if (!due(f, level)) continue;
            vct_copy_alpha_desc d{(int32_t)level, (int32_t)_cfg.voxel_resolution, 6};
            VCT_PASS_CHECK(vct_copy_alpha(_ctx, f->stream, &d));
        }
        for (uint32_t level = 1; level < _cfg.clip_levels; ++level) {
            if (!due(f, level)) continue;
            const vct_clipmap_region prev = clip_region(_ctx, level - 1, _cfg.clip_levels);
            vct_downsample_desc d{{prev.min_corner[0], prev.min_corner[1], prev.min_corner[2]},
                                  (int32_t)level,
                                  (int32_t)_cfg.voxel_resolution,
                                  (int32_t)_cfg.downsample_region_size};
            VCT_PASS_CHECK(vct_downsample(_ctx, f->stream, VCT_VOLUME_RADIANCE, &d));
        }
    }

private:
    vct_ctx* _ctx;
    vct_config _cfg;
    const vct_mesh* _mesh;
    const vct_shadow_map* _shadow;
    const vct_directional_light_desc* _light;
    const vct_directional_light_shadow_desc* _light_shadow;
};

// VoxelConeTracingPass::onUpdate (VoxelConeTracingPass.cpp:75-106) with the
// member defaults of VoxelConeTracingPass.h:75-82; the volume centre and
// voxel size come from clip level 0 as the pass computes them (:85-92).
class VoxelConeTracingPassHip : public RenderPassBase {
public:
    VoxelConeTracingPassHip(vct_ctx* ctx, const vct_config& cfg, const vct_gbuffer* gbuffer,
                            const vct_camera_ubo* camera, const vct_shadow_map* shadow,
                            const vct_directional_light_desc* light,
                            const vct_directional_light_shadow_desc* light_shadow, float* diffuse_out,
                            float* specular_out)
        : _ctx(ctx), _cfg(cfg), _gbuffer(gbuffer), _camera(camera), _shadow(shadow), _light(light),
          _light_shadow(light_shadow), _diffuse(diffuse_out), _specular(specular_out)
    {
    }

    uint32_t rendering_mode = VCT_MODE_COMBINED_GI;
    float trace_start_offset = 1.0f;
    float indirect_diffuse_intensity = 8.0f;
    float ambient_occlusion_factor = 0.5f;
    float min_trace_step_factor = 1.0f;
    float indirect_specular_intensity = 3.0f;
    float occlusion_decay = 2.0f;
    bool enable_32_cones = false;
    unsigned long long* samples = nullptr;  // optional device counter

    vct_cone_tracing_desc desc() const
    {
        const vct_clipmap_region r = clip_region(_ctx, 0, _cfg.clip_levels);
        vct_cone_tracing_desc d{};
        for (int k = 0; k < 3; ++k) {
            const float lo = (float)r.min_corner[k] * r.voxel_size;
            const float half = ((float)r.extent[k] * r.voxel_size) * 0.5f;
            d.volume_center[k] = lo + half;
        }
        d.rendering_mode = rendering_mode;
        d.voxel_size = r.voxel_size;
        d.volume_dimension = (float)_cfg.voxel_resolution;
        d.trace_start_offset = trace_start_offset;
        d.indirect_diffuse_intensity = indirect_diffuse_intensity;
        d.ambient_occlusion_factor = ambient_occlusion_factor;
        d.min_trace_step_factor = min_trace_step_factor;
        d.indirect_specular_intensity = indirect_specular_intensity;
        d.occlusion_decay = occlusion_decay;
        d.enable_32_cones = enable_32_cones ? 1 : 0;
        return d;
    }

protected:
    void onUpdate(const FrameLayout* f) override
    {
        const vct_cone_tracing_desc d = desc();
        VCT_PASS_CHECK(vct_trace(_ctx, f->stream, _gbuffer, _camera, &d, _shadow, _light, _light_shadow, _diffuse,
                                 _specular, samples));
    }

private:
    vct_ctx* _ctx;
    vct_config _cfg;
    const vct_gbuffer* _gbuffer;
    const vct_camera_ubo* _camera;
    const vct_shadow_map* _shadow;
    const vct_directional_light_desc* _light;
    const vct_directional_light_shadow_desc* _light_shadow;
    float* _diffuse;
    float* _specular;
};

}  // namespace vct
