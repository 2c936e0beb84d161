// vct_torch.cpp - libtorch binding of the dense method's forward render and
// its hand-written adjoint as a torch::autograd::Function.
//
// render(albedo[R,R,R,3], opacity[R,R,R], diffuse[V,H,W] int32,
//        normal[V,H,W,4] int16, depth[V,H,W] float, cameras[V,36] float,
//        grid = (min x, min y, min z, extent), trace = (aperture, max distance,
//        step factor, start offset, occlusion decay, AO factor, indirect
//        intensity, 32 cones), flags) -> out[V,H,W,4]
//
// The forward builds the level-0 texels, the anisotropic mips and traces
// every view (vct_dense_set_level0 / vct_dense_mipmap / vct_dense_trace).
// The backward reuses that pyramid when the grid still holds it (an id per
// build, recorded by the forward; else it builds it again from the saved
// inputs; the grid object is shared between calls),
// zeroes a gradient pyramid, runs the trace adjoint (from the forward's tapes
// in one vct_dense_trace_backward_tapes call, or vct_dense_trace_backward
// per view), vct_dense_mipmap_backward, and returns d loss / d albedo and d loss /
// d opacity. All work is enqueued on the current HIP stream of PyTorch, and
// the backward does not wait for it: a trace status set on the device turns
// the gradients' first element into NaN there, and the next backward (or
// check_status()) raises on it. `cameras` rows are view_proj (16),
// view_proj_inv (16), eye (3), padding (1), column-major as glm stores them.
#include <torch/extension.h>
#include <ATen/hip/HIPContext.h>

#include <array>
#include <cmath>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/vct.h"

namespace {

void check(vct_status s, const char* what)
{
    if (s != VCT_OK) throw std::runtime_error(std::string(what) + ": " + vct_last_error());
}

// One RGBA32F grid per (device, resolution, box), reused across calls.
struct GridKey {
    int device;
    int64_t res;
    std::array<float, 4> box;
    bool operator==(const GridKey& o) const { return device == o.device && res == o.res && box == o.box; }
};

struct GridSlot {
    GridKey key{};
    vct_dense* grid = nullptr;
    size_t texels = 0;
    // id of the pyramid the grid holds (0: none); every build takes a new one
    int64_t build = 0;
};

std::mutex g_mu;
std::deque<GridSlot> g_grids;  // stable addresses: slots are handed out by reference
int64_t g_builds = 0;          // pyramid ids handed out

// The trace backward's status word of the previous backward, copied to pinned
// memory without waiting; checked at the start of the next backward.
struct PendingStatus {
    int32_t* host = nullptr;
    hipEvent_t done = nullptr;
    bool pending = false;
};
PendingStatus g_status;

GridSlot& grid_for(int device, int64_t res, const std::array<float, 4>& box)
{
    std::lock_guard<std::mutex> lock(g_mu);
    GridKey k{device, res, box};
    for (auto& s : g_grids)
        if (s.key == k) return s;
    GridSlot s;
    s.key = k;
    vct_dense_desc d{};
    d.grid_min[0] = box[0];
    d.grid_min[1] = box[1];
    d.grid_min[2] = box[2];
    d.extent = box[3];
    d.resolution = (uint32_t)res;
    d.format = VCT_FORMAT_RGBA32F;
    d.raster_samples = 1;
    check(vct_dense_create(&s.grid, &d), "vct_dense_create");
    check(vct_dense_texels(s.grid, &s.texels), "vct_dense_texels");
    g_grids.push_back(s);
    return g_grids.back();
}

vct_dense_trace_desc trace_desc(const std::vector<double>& t)
{
    TORCH_CHECK(t.size() == 8, "trace parameters: 8 values");
    vct_dense_trace_desc d{};
    d.aperture = (float)t[0];
    d.max_distance = (float)t[1];
    d.min_trace_step_factor = (float)t[2];
    d.trace_start_offset = (float)t[3];
    d.occlusion_decay = (float)t[4];
    d.ambient_occlusion_factor = (float)t[5];
    d.indirect_diffuse_intensity = (float)t[6];
    d.enable_32_cones = t[7] != 0.0 ? 1 : 0;
    return d;
}

vct_camera_ubo camera(const torch::Tensor& cams, int64_t v)
{
    vct_camera_ubo c{};
    const float* p = cams.data_ptr<float>() + 36 * v;
    std::memcpy(c.view_proj, p, 16 * sizeof(float));
    std::memcpy(c.view_proj_inv, p + 16, 16 * sizeof(float));
    std::memcpy(c.eye_pos, p + 32, 3 * sizeof(float));
    return c;
}

vct_gbuffer gbuffer(const torch::Tensor& diffuse, const torch::Tensor& normal, const torch::Tensor& depth, int64_t v)
{
    vct_gbuffer g{};
    const int64_t H = depth.size(1), W = depth.size(2);
    g.diffuse = (const uint32_t*)(diffuse.data_ptr<int32_t>() + v * H * W);
    g.normal = (const uint16_t*)(normal.data_ptr<int16_t>() + v * H * W * 4);
    g.depth = depth.data_ptr<float>() + v * H * W;
    g.specular = g.diffuse;  // unused by the dense trace
    g.emission = g.normal;   // unused by the dense trace
    g.width = (uint32_t)W;
    g.height = (uint32_t)H;
    return g;
}

void check_inputs(const torch::Tensor& albedo, const torch::Tensor& opacity, const torch::Tensor& diffuse,
                  const torch::Tensor& normal, const torch::Tensor& depth, const torch::Tensor& cams)
{
    TORCH_CHECK(albedo.is_cuda() && albedo.scalar_type() == torch::kFloat32 && albedo.dim() == 4 &&
                    albedo.size(3) == 3 && albedo.size(0) == albedo.size(1) && albedo.size(1) == albedo.size(2),
                "albedo: float32 device tensor [R, R, R, 3]");
    const int64_t R = albedo.size(0);
    TORCH_CHECK(opacity.is_cuda() && opacity.scalar_type() == torch::kFloat32 && opacity.dim() == 3 &&
                    opacity.size(0) == R && opacity.size(1) == R && opacity.size(2) == R,
                "opacity: float32 device tensor [R, R, R]");
    TORCH_CHECK(depth.is_cuda() && depth.scalar_type() == torch::kFloat32 && depth.dim() == 3 && depth.is_contiguous(),
                "depth: contiguous float32 device tensor [V, H, W]");
    TORCH_CHECK(diffuse.is_cuda() && diffuse.scalar_type() == torch::kInt32 && diffuse.sizes() == depth.sizes() &&
                    diffuse.is_contiguous(),
                "diffuse: contiguous int32 device tensor [V, H, W]");
    TORCH_CHECK(normal.is_cuda() && normal.scalar_type() == torch::kInt16 && normal.dim() == 4 &&
                    normal.size(0) == depth.size(0) && normal.size(1) == depth.size(1) &&
                    normal.size(2) == depth.size(2) && normal.size(3) == 4 && normal.is_contiguous(),
                "normal: contiguous int16 (half bits) device tensor [V, H, W, 4]");
    TORCH_CHECK(!cams.is_cuda() && cams.scalar_type() == torch::kFloat32 && cams.dim() == 2 && cams.size(1) == 36 &&
                    cams.size(0) == depth.size(0) && cams.is_contiguous(),
                "cameras: contiguous float32 host tensor [V, 36]");
}

// Level 0 and the mips of `albedo`/`opacity` in the shared grid. A grid that
// still holds pyramid `reuse` (the id the forward recorded; autograd refuses
// a backward whose saved inputs changed in place) is used as it is: the
// backward reuses the forward's pyramid unless another render ran in between.
GridSlot& build_pyramid(const torch::Tensor& albedo, const torch::Tensor& opacity, const std::array<float, 4>& box,
                        hipStream_t s, int64_t reuse)
{
    GridSlot& slot = grid_for(albedo.get_device(), albedo.size(0), box);
    if (reuse != 0 && slot.build == reuse) return slot;
    check(vct_dense_set_level0(slot.grid, s, albedo.data_ptr<float>(), opacity.data_ptr<float>()),
          "vct_dense_set_level0");
    check(vct_dense_mipmap(slot.grid, s), "vct_dense_mipmap");
    std::lock_guard<std::mutex> lock(g_mu);
    slot.build = ++g_builds;
    return slot;
}

// Raises if the previous backward's trace status was set (its gradient was
// already poisoned with NaN on the device).
void check_previous_status()
{
    if (!g_status.pending) return;
    g_status.pending = false;
    check(hipEventSynchronize(g_status.done) == hipSuccess ? VCT_OK : VCT_ERR_HIP, "status event");
    const int32_t st = *g_status.host;
    TORCH_CHECK(st == 0, "a previous vct_dense_trace_backward reported status ", st,
                " (bit 0: a march outgrew its per-lane state; bit 1: the gradient records outgrew their buffer); "
                "its gradient holds NaN");
}

// Copies the status word to pinned memory behind the backward's work and
// poisons the gradients with NaN on the device when it is set (no host wait).
void defer_status(const torch::Tensor& status, torch::Tensor& ga, torch::Tensor& go, hipStream_t s)
{
    if (!g_status.host) {
        check(hipHostMalloc((void**)&g_status.host, sizeof(int32_t), hipHostMallocDefault) == hipSuccess
                  ? VCT_OK : VCT_ERR_OUT_OF_MEMORY, "pinned status word");
        check(hipEventCreateWithFlags(&g_status.done, hipEventDisableTiming) == hipSuccess ? VCT_OK : VCT_ERR_HIP,
              "status event");
    }
    const auto bad = status.ne(0);
    ga.view(-1).narrow(0, 0, 1).masked_fill_(bad, std::nanf(""));
    go.view(-1).narrow(0, 0, 1).masked_fill_(bad, std::nanf(""));
    check(hipMemcpyAsync(g_status.host, status.data_ptr<int32_t>(), sizeof(int32_t), hipMemcpyDeviceToHost, s) ==
              hipSuccess ? VCT_OK : VCT_ERR_HIP, "status copy");
    check(hipEventRecord(g_status.done, s) == hipSuccess ? VCT_OK : VCT_ERR_HIP, "status event");
    g_status.pending = true;
}

std::array<float, 4> box_of(const std::vector<double>& grid)
{
    TORCH_CHECK(grid.size() == 4, "grid: (min x, min y, min z, extent)");
    return {(float)grid[0], (float)grid[1], (float)grid[2], (float)grid[3]};
}

// Flag bits of render() beside VCT_ADJOINT_RECOMPUTE: kNoTape (the caller's)
// makes the backward march again with the samples kept per lane instead of
// reading the forward's tape; kTapeWanted (set by render) says a backward
// can follow, so the forward may keep the tape.
constexpr int64_t kNoTape = int64_t(1) << 1;
constexpr int64_t kTapeWanted = int64_t(1) << 32;

class DenseRender : public torch::autograd::Function<DenseRender> {
public:
    static torch::Tensor forward(torch::autograd::AutogradContext* ctx, torch::Tensor albedo, torch::Tensor opacity,
                                 torch::Tensor diffuse, torch::Tensor normal, torch::Tensor depth,
                                 torch::Tensor cameras, std::vector<double> grid, std::vector<double> trace,
                                 int64_t flags)
    {
        check_inputs(albedo, opacity, diffuse, normal, depth, cameras);
        albedo = albedo.contiguous();
        opacity = opacity.contiguous();
        const auto box = box_of(grid);
        const vct_dense_trace_desc td = trace_desc(trace);
        hipStream_t s = c10::hip::getCurrentHIPStream().stream();
        GridSlot& slot = build_pyramid(albedo, opacity, box, s, 0);
        const int64_t V = depth.size(0), H = depth.size(1), W = depth.size(2);
        auto out = torch::empty({V, H, W, 4}, albedo.options());
        // When a gradient will be asked for, the forward keeps the march's
        // samples (the adjoint tape) unless told to recompute them or they
        // would take more than half of the free device memory.
        const bool grad = (flags & kTapeWanted) != 0;
        torch::Tensor tape;
        int64_t per_view = 0;
        if (grad && !(flags & (VCT_ADJOINT_RECOMPUTE | kNoTape)) && V > 0) {
            const vct_gbuffer gb = gbuffer(diffuse, normal, depth, 0);
            const vct_camera_ubo cam = camera(cameras, 0);
            size_t bytes = 0;
            size_t free_b = 0, total_b = 0;
            if (vct_dense_tape_bytes(slot.grid, &gb, &cam, &td, &bytes) == VCT_OK &&
                hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
                per_view = (int64_t)((bytes + 255) / 256 * 256);
                if ((size_t)(per_view * V) <= free_b / 2)
                    tape = torch::empty({per_view * V}, albedo.options().dtype(torch::kUInt8));
            }
        }
        for (int64_t v = 0; v < V; ++v) {
            const vct_gbuffer gb = gbuffer(diffuse, normal, depth, v);
            const vct_camera_ubo cam = camera(cameras, v);
            float* o = out.data_ptr<float>() + v * H * W * 4;
            if (tape.defined())
                check(vct_dense_trace_tape(slot.grid, s, &gb, &cam, &td, o, nullptr,
                                           tape.data_ptr<uint8_t>() + v * per_view, (size_t)per_view),
                      "vct_dense_trace_tape");
            else
                check(vct_dense_trace(slot.grid, s, &gb, &cam, &td, o, nullptr), "vct_dense_trace");
        }
        if (tape.defined()) {
            ctx->saved_data["tape"] = tape;
            ctx->saved_data["tape_view_bytes"] = per_view;
        }
        ctx->save_for_backward({albedo, opacity, diffuse, normal, depth, cameras});
        ctx->saved_data["grid"] = grid;
        ctx->saved_data["trace"] = trace;
        ctx->saved_data["flags"] = flags;
        ctx->saved_data["build"] = slot.build;
        return out;
    }

    static torch::autograd::tensor_list backward(torch::autograd::AutogradContext* ctx,
                                                 torch::autograd::tensor_list grads)
    {
        const auto saved = ctx->get_saved_variables();
        const torch::Tensor &albedo = saved[0], &opacity = saved[1], &diffuse = saved[2], &normal = saved[3],
                            &depth = saved[4], &cameras = saved[5];
        const auto box = box_of(ctx->saved_data["grid"].toDoubleVector());
        const vct_dense_trace_desc td = trace_desc(ctx->saved_data["trace"].toDoubleVector());
        const uint32_t flags = (uint32_t)(ctx->saved_data["flags"].toInt() & VCT_ADJOINT_RECOMPUTE);
        hipStream_t s = c10::hip::getCurrentHIPStream().stream();
        check_previous_status();
        GridSlot& slot = build_pyramid(albedo, opacity, box, s, ctx->saved_data["build"].toInt());
        torch::Tensor gout = grads[0].contiguous();
        auto gpyr = torch::zeros({(int64_t)slot.texels * 4}, albedo.options());
        auto status = torch::zeros({1}, albedo.options().dtype(torch::kInt32));
        const int64_t V = depth.size(0), H = depth.size(1), W = depth.size(2);
        // The tape serves one backward (it is overwritten with the sample
        // gradients); a second backward of a retained graph marches again.
        torch::Tensor tape;
        if (ctx->saved_data.count("tape")) {
            tape = ctx->saved_data["tape"].toTensor();
            ctx->saved_data.erase("tape");
        }
        const int64_t per_view = tape.defined() ? ctx->saved_data["tape_view_bytes"].toInt() : 0;
        auto* st = (uint32_t*)status.data_ptr<int32_t>();
        if (tape.defined()) {  // all views in one call: they share the launches
            std::vector<vct_gbuffer> gbs;
            std::vector<vct_camera_ubo> cams;
            std::vector<const float*> gouts;
            std::vector<void*> tapes;
            for (int64_t v = 0; v < V; ++v) {
                gbs.push_back(gbuffer(diffuse, normal, depth, v));
                cams.push_back(camera(cameras, v));
                gouts.push_back(gout.data_ptr<float>() + v * H * W * 4);
                tapes.push_back(tape.data_ptr<uint8_t>() + v * per_view);
            }
            check(vct_dense_trace_backward_tapes(slot.grid, s, (uint32_t)V, gbs.data(), cams.data(), &td, gouts.data(),
                                                 tapes.data(), (size_t)per_view, gpyr.data_ptr<float>(), st),
                  "vct_dense_trace_backward_tapes");
        } else {
            for (int64_t v = 0; v < V; ++v) {
                const vct_gbuffer gb = gbuffer(diffuse, normal, depth, v);
                const vct_camera_ubo cam = camera(cameras, v);
                check(vct_dense_trace_backward(slot.grid, s, &gb, &cam, &td, gout.data_ptr<float>() + v * H * W * 4,
                                               gpyr.data_ptr<float>(), flags, st),
                      "vct_dense_trace_backward");
            }
        }
        check(vct_dense_mipmap_backward(slot.grid, s, gpyr.data_ptr<float>()), "vct_dense_mipmap_backward");
        auto ga = torch::empty_like(albedo);
        auto go = torch::empty_like(opacity);
        check(vct_dense_level0_grad(slot.grid, s, gpyr.data_ptr<float>(), ga.data_ptr<float>(), go.data_ptr<float>()),
              "vct_dense_level0_grad");
        defer_status(status, ga, go, s);
        return {ga, go, torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor(),
                torch::Tensor(), torch::Tensor()};
    }
};

torch::Tensor render(torch::Tensor albedo, torch::Tensor opacity, torch::Tensor diffuse, torch::Tensor normal,
                     torch::Tensor depth, torch::Tensor cameras, std::vector<double> grid, std::vector<double> trace,
                     int64_t flags)
{
    // what apply() will find: a backward can follow only in grad mode with an
    // input that requires a gradient
    if (torch::GradMode::is_enabled() && (albedo.requires_grad() || opacity.requires_grad())) flags |= kTapeWanted;
    return DenseRender::apply(albedo, opacity, diffuse, normal, depth, cameras, grid, trace, flags);
}

void release_grids()
{
    std::lock_guard<std::mutex> lock(g_mu);
    for (auto& s : g_grids) vct_dense_destroy(s.grid);
    g_grids.clear();
}

// Waits for the last backward and raises if its trace status was set.
void check_status() { check_previous_status(); }

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m)
{
    m.doc() = "dense-method render with a hand-written adjoint (torch.autograd.Function over libvct)";
    m.def("render", &render, "differentiable dense VXAO / indirect diffuse render");
    m.def("release_grids", &release_grids, "frees the cached device grids");
    m.def("check_status", &check_status, "raises if the last backward's trace status was set");
}
