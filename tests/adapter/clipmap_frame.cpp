// clipmap_frame.cpp - drives one clipmap GI frame through the C++ pass
// adapter (include/vct_passes.hpp) from a compiled translation unit: the
// inputs come from a record file written by the test, the outputs (both
// volumes in the reference atlas layout and the two trace images) go to
// another record file.
//
// Record file: repeated {u32 name length, name, u64 byte count, bytes}.
// Inputs: config, eye, positions, normals, texcoords, indices, nodes,
// materials, primitives, bb_min, bb_max, texture_count and texture<i> +
// texture<i>_dims (RGBA8 texels, width and height), shadow (floats) + shadow_dims,
// light, light_shadow, camera, gbuffer_{diffuse,normal,specular,emission,depth}
// + gbuffer_dims.
#include <hip/hip_runtime_api.h>

#include <array>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "../../include/vct_passes.hpp"

using Blob = std::vector<unsigned char>;

static std::map<std::string, Blob> read_records(const char* path)
{
    std::map<std::string, Blob> out;
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error(std::string("cannot open ") + path);
    uint32_t n = 0;
    while (in.read((char*)&n, 4)) {
        std::string name(n, '\0');
        uint64_t bytes = 0;
        in.read(&name[0], n);
        in.read((char*)&bytes, 8);
        Blob b(bytes);
        if (bytes) in.read((char*)b.data(), (std::streamsize)bytes);
        if (!in) throw std::runtime_error("truncated record " + name);
        out[name] = std::move(b);
    }
    return out;
}

static void write_record(std::ofstream& o, const std::string& name, const void* p, uint64_t bytes)
{
    const uint32_t n = (uint32_t)name.size();
    o.write((const char*)&n, 4);
    o.write(name.data(), n);
    o.write((const char*)&bytes, 8);
    o.write((const char*)p, (std::streamsize)bytes);
}

template <typename T>
static T pod(const std::map<std::string, Blob>& r, const char* name)
{
    const Blob& b = r.at(name);
    if (b.size() != sizeof(T)) throw std::runtime_error(std::string("size of record ") + name);
    T v;
    std::memcpy(&v, b.data(), sizeof(T));
    return v;
}

#define HIP_CHECK(expr)                                                                              \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct DeviceCopies {
    std::vector<void*> ptrs;
    void* up(const Blob& b)
    {
        void* p = nullptr;
        HIP_CHECK(hipMalloc(&p, b.size() ? b.size() : 4));
        if (b.size()) HIP_CHECK(hipMemcpy(p, b.data(), b.size(), hipMemcpyHostToDevice));
        ptrs.push_back(p);
        return p;
    }
    void* zeros(size_t bytes)
    {
        void* p = nullptr;
        HIP_CHECK(hipMalloc(&p, bytes));
        HIP_CHECK(hipMemset(p, 0, bytes));
        ptrs.push_back(p);
        return p;
    }
    ~DeviceCopies()
    {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s inputs.rec outputs.rec\n", argv[0]);
        return 2;
    }
    try {
        const auto rec = read_records(argv[1]);
        const vct_config cfg = pod<vct_config>(rec, "config");
        const auto eye = pod<std::array<float, 3>>(rec, "eye");
        DeviceCopies dev;
        vct_mesh mesh{};
        mesh.positions = (const float*)dev.up(rec.at("positions"));
        mesh.normals = (const float*)dev.up(rec.at("normals"));
        mesh.texcoords = (const float*)dev.up(rec.at("texcoords"));
        mesh.indices = (const uint32_t*)dev.up(rec.at("indices"));
        mesh.node_matrices = (const float*)dev.up(rec.at("nodes"));
        mesh.materials = (const vct_gltf_material*)dev.up(rec.at("materials"));
        mesh.vertex_count = (uint32_t)(rec.at("positions").size() / 12);
        mesh.index_count = (uint32_t)(rec.at("indices").size() / 4);
        mesh.instance_count = (uint32_t)(rec.at("nodes").size() / 128);
        mesh.material_count = (uint32_t)(rec.at("materials").size() / sizeof(vct_gltf_material));
        const uint32_t ntex = pod<uint32_t>(rec, "texture_count");
        std::vector<vct_texture> textures(ntex);
        for (uint32_t i = 0; i < ntex; ++i) {
            const std::string name = "texture" + std::to_string(i);
            const auto dims = pod<std::array<uint32_t, 2>>(rec, (name + "_dims").c_str());
            if (rec.at(name).size() != (size_t)dims[0] * dims[1] * 4u) throw std::runtime_error("size of " + name);
            textures[i].texels = (const uint32_t*)dev.up(rec.at(name));
            textures[i].width = dims[0];
            textures[i].height = dims[1];
        }
        mesh.textures = ntex ? textures.data() : nullptr;
        mesh.texture_count = ntex;
        const Blob& prims = rec.at("primitives");
        mesh.primitives = (const vct_primitive*)prims.data();
        mesh.primitive_count = (uint32_t)(prims.size() / sizeof(vct_primitive));
        const auto bmin = pod<std::array<float, 3>>(rec, "bb_min");
        const auto bmax = pod<std::array<float, 3>>(rec, "bb_max");
        for (int k = 0; k < 3; ++k) {
            mesh.scene_bb_min[k] = bmin[k];
            mesh.scene_bb_max[k] = bmax[k];
        }
        const auto sdims = pod<std::array<uint32_t, 2>>(rec, "shadow_dims");
        const vct_shadow_map shadow{(const float*)dev.up(rec.at("shadow")), sdims[0], sdims[1]};
        const auto light = pod<vct_directional_light_desc>(rec, "light");
        const auto light_shadow = pod<vct_directional_light_shadow_desc>(rec, "light_shadow");
        const auto camera = pod<vct_camera_ubo>(rec, "camera");
        const auto gdims = pod<std::array<uint32_t, 2>>(rec, "gbuffer_dims");
        const vct_gbuffer gb{(const uint32_t*)dev.up(rec.at("gbuffer_diffuse")),
                             (const uint16_t*)dev.up(rec.at("gbuffer_normal")),
                             (const uint32_t*)dev.up(rec.at("gbuffer_specular")),
                             (const uint16_t*)dev.up(rec.at("gbuffer_emission")),
                             (const float*)dev.up(rec.at("gbuffer_depth")), gdims[0], gdims[1]};
        const size_t img_bytes = (size_t)gdims[0] * gdims[1] * 16u;
        float* diffuse = (float*)dev.zeros(img_bytes);
        float* specular = (float*)dev.zeros(img_bytes);

        vct_ctx* ctx = nullptr;
        VCT_PASS_CHECK(vct_create(&ctx, &cfg));
        hipStream_t stream = nullptr;
        HIP_CHECK(hipStreamCreate(&stream));
        {
            VCT_PASS_CHECK(vct_update_clip_regions(ctx, eye.data()));
            vct::RenderPassManagerHip mgr;
            mgr.addRenderPass<vct::VoxelizationPassHip>("VoxelizationPass", ctx, cfg, &mesh);
            mgr.addRenderPass<vct::RadianceInjectionPassHip>("RadianceInjectionPass", ctx, cfg, &mesh, &shadow,
                                                              &light, &light_shadow);
            mgr.addRenderPass<vct::VoxelConeTracingPassHip>("VoxelConeTracingPass", ctx, cfg, &gb, &camera, &shadow,
                                                             &light, &light_shadow, diffuse, specular);
            vct::FrameLayout frame;
            frame.stream = stream;
            for (const char* name : {"VoxelizationPass", "RadianceInjectionPass", "VoxelConeTracingPass"})
                mgr.drawSingleRenderPass(name, &frame);
            HIP_CHECK(hipStreamSynchronize(stream));

            std::ofstream out(argv[2], std::ios::binary);
            for (vct_volume v : {VCT_VOLUME_OPACITY, VCT_VOLUME_RADIANCE}) {
                size_t bytes = 0;
                VCT_PASS_CHECK(vct_volume_bytes(ctx, v, &bytes));
                Blob host(bytes);
                VCT_PASS_CHECK(vct_read_volume(ctx, stream, v, host.data(), bytes));
                HIP_CHECK(hipStreamSynchronize(stream));
                write_record(out, v == VCT_VOLUME_OPACITY ? "opacity" : "radiance", host.data(), bytes);
            }
            for (auto [name, p] : {std::pair<const char*, float*>{"diffuse", diffuse}, {"specular", specular}}) {
                Blob host(img_bytes);
                HIP_CHECK(hipMemcpy(host.data(), p, img_bytes, hipMemcpyDeviceToHost));
                write_record(out, name, host.data(), img_bytes);
            }
        }
        HIP_CHECK(hipStreamDestroy(stream));
        VCT_PASS_CHECK(vct_destroy(ctx));
        std::printf("clipmap_frame: %u x %u, %u levels of %u^3: done\n", gdims[0], gdims[1], cfg.clip_levels,
                    cfg.voxel_resolution);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "clipmap_frame: %s\n", e.what());
        return 1;
    }
}
