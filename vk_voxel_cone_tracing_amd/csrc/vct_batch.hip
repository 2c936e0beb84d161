// vct_batch.hip - the batched dense method: many scenes of uneven size
// voxelized, mipped and traced in ONE persistent launch, replayed from a
// hipGraph.
//
// The launch is a pool of resident workgroups pulling tasks from a queue
// (an atomic head over a task table built on the host). Tasks are ordered by
// phase - vertex chunks, raster chunks, mip bricks level by level, pixel
// tiles - and every task of a phase depends only on its own scene's previous
// phase, tracked by one completion counter per (scene, phase). A workgroup
// takes a task only while it runs, and every task waits only on tasks earlier
// in the queue, which were taken by running workgroups, so the waits always
// end; a bounded spin with a shared abort flag guarantees that every wave
// exits even if that reasoning were broken. Small scenes reach their trace
// tiles while large scenes are still voxelizing, and a scene's tiny mip
// levels never leave the GPU idle.
//
// The phased schedule (workgroups_per_cu = 0) runs the same work as one flat
// launch per phase instead - vertex chunks; the raster of every scene's
// triangles as one count / scan / sample-parallel pass (the per-scene
// voxelizer's scheme over the whole batch); each mip level; pixel tiles -
// each kernel sized to its own task type, with kernel boundaries as the only
// dependencies; the graph replays them back to back.
//
// Each task runs the same device code as the per-scene calls
// (vct_dense_voxelize / vct_dense_mipmap / vct_dense_trace), so the grids are
// bit-identical to theirs and the traced images equal theirs.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "vct_dense.hpp"
#include "vct_dense_march.hpp"
#include "vct_dense_mip.hpp"
#include "vct_internal.hpp"
#include "vct_raster.hpp"
#include "vct_texel.hpp"

static_assert(sizeof(vct_batch_tuning) == 16, "vct_batch_tuning layout");
static_assert(sizeof(vct_batch_scene) == 352, "vct_batch_scene layout");
static_assert(sizeof(vct_batch_scene_shape) == 16, "vct_batch_scene_shape layout");

namespace vct {
namespace batch {

enum TaskType : uint32_t { kVertex = 0, kRaster = 1, kMip = 2, kTrace = 3 };

// Phase p of a scene: 0 vertex, 1 raster, 1 + m mip level m (m >= 1).
constexpr int kPhases = 1 + dense::kMaxLevels;
static_assert(kPhases == VCT_BATCH_PHASES, "phases per scene");
constexpr int kTile = 16;                    // trace tile edge (pixels)
constexpr double kWaitSeconds = 2.0;         // a dependency wait gives up (and aborts the launch) after this

struct Task {
    uint32_t type_scene;  // type << 24 | scene
    uint32_t a;           // first triangle / brick index / tile index
    uint32_t b;           // end triangle (vertex, raster)
    uint32_t level;       // mip level (kMip)
};

struct Prim {
    uint32_t first_index, tri_count;
    int32_t vertex_offset;
    uint32_t instance_index, material_index, first_tri;
};

struct SceneDev {
    const float* positions;
    const uint32_t* indices;
    const float* nodes;
    const vct_gltf_material* materials;
    const Prim* prims;
    uint32_t nprims;
    uint32_t tri_count;
    float* soup_pos;     // 9 floats per triangle, world space
    uint32_t* soup_mat;  // material per triangle
    float grid_min[3];
    float vs;
    int res;
    int samples;  // raster samples
    uint32_t* winner;
    void* data;  // pyramid base (RGBA8)
    dense::Pyramid pyr;
    dense::March march;
    const uint32_t* g_diffuse;
    const uint16_t* g_normal;
    const float* g_depth;
    int width, height, tiles_x;
    float4* out;
    uint32_t need[kPhases];  // tasks of each phase
};

struct Args {
    const Task* tasks;
    uint32_t ntasks;
    uint32_t first;  // the persistent launch's first task (the first mip brick)
    const SceneDev* scenes;
    uint32_t* head;
    uint32_t* done;  // [scene * kPhases + phase]
    uint32_t* abort_flag;
    uint32_t* status;
    unsigned long long* samples;
    uint32_t mip_brick;
    uint32_t schedule;
    long long wait_ticks;  // kWaitSeconds in wall-clock ticks
};

__global__ void k_batch_reset(uint32_t* head, uint32_t* done, uint32_t ndone, uint32_t* abort_flag)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ndone) done[i] = 0u;
    if (i == 0) {
        *head = 0u;
        *abort_flag = 0u;
    }
}

// Thread 0 polls the counter; false if the launch is aborting. The wait is
// bounded by wall-clock time, whatever one poll costs under contention.
__device__ bool wait_count(const uint32_t* counter, uint32_t need, uint32_t* abort_flag, uint32_t* status,
                           long long ticks)
{
    const long long t0 = wall_clock64();
    while (__hip_atomic_load(counter, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < need) {
        if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
        __builtin_amdgcn_s_sleep(2);
        if (wall_clock64() - t0 > ticks) {
            __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            atomicOr(status, 1u);
            return false;
        }
    }
    return true;
}

// msaaVoxelizer.vert:31-37 for triangles [a, b) of one scene (one lane each).
__device__ void vertex_task(const SceneDev& S, uint32_t a, uint32_t b)
{
#pragma clang fp contract(off)
    const uint32_t t = a + threadIdx.x;
    if (t >= b) return;
    uint32_t lo = 0, hi = S.nprims;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (S.prims[mid].first_tri <= t) lo = mid; else hi = mid;
    }
    const Prim pr = S.prims[lo];
    const uint32_t local = t - pr.first_tri;
    const float* m = S.nodes + 32u * pr.instance_index;
    for (int k = 0; k < 3; ++k) {
        const uint32_t vi = (uint32_t)((int64_t)S.indices[pr.first_index + 3u * local + (uint32_t)k] + pr.vertex_offset);
        const float px = S.positions[3u * vi], py = S.positions[3u * vi + 1], pz = S.positions[3u * vi + 2];
        for (int r = 0; r < 3; ++r)
            S.soup_pos[9u * t + 3 * k + r] = ((m[r] * px + m[4 + r] * py) + m[8 + r] * pz) + m[12 + r];
    }
    S.soup_mat[t] = pr.material_index;
}

// One brick of mip level m. Level 1 resolves its 8 level-0 children from the
// winners (k_dense_resolve), writes them to level 0 and returns the winners
// to empty for the next launch. kLean (the persistent launch) keeps few
// registers live: one face at a time, and level 1 reads its children back
// from level 0 after writing them, so that the pixel tiles sharing the
// launch keep four waves per SIMD; the flat mip launch keeps every load of a
// texel in flight at once instead.
template <bool kLean>
__device__ void mip_task(const SceneDev& S, uint32_t brick, int m, uint32_t B)
{
    const int n = S.res >> m;
    const uint32_t be = (uint32_t)(n < (int)B ? n : (int)B);
    const uint32_t nb = ((uint32_t)n + be - 1) / be;
    const uint32_t bx = brick % nb, by = (brick / nb) % nb, bz = brick / (nb * nb);
    const uint32_t cells = be * be * be;
    char* base = (char*)S.data;
    const void* src = base + 4u * S.pyr.offset[m - 1];
    void* dst = base + 4u * S.pyr.offset[m];
    for (uint32_t i = threadIdx.x; i < cells; i += blockDim.x) {
        const int x = (int)(bx * be + i % be), y = (int)(by * be + (i / be) % be), z = (int)(bz * be + i / (be * be));
        if (x >= n || y >= n || z >= n) continue;
        if (m > 1) {
            if (kLean) dense_mip_texel_lean<false>(src, dst, n, x, y, z, false);
            else dense_mip_texel<false>(src, dst, n, x, y, z, false);
            continue;
        }
        const size_t D0 = (size_t)S.res + 2;
        auto resolve = [&](int c) -> uint32_t {
            const int cx = 2 * x + (c & 1), cy = 2 * y + ((c >> 1) & 1), cz = 2 * z + (c >> 2);
            uint32_t* wp = &S.winner[((size_t)cz * S.res + cy) * S.res + cx];
            const uint32_t t = *wp;
            uint32_t packed = 0u;
            if (t != 0xffffffffu) {
                const float* bc = S.materials[S.soup_mat[t]].base_color_factor;
                packed = pack_unorm(f4{bc[0], bc[1], bc[2], 1.0f});
                *wp = 0xffffffffu;
            }
            ((uint32_t*)src)[((size_t)(cz + 1) * D0 + (size_t)(cy + 1)) * D0 + (size_t)(cx + 1)] = packed;
            return packed;
        };
        if (kLean) {
#pragma unroll 1
            for (int c = 0; c < 8; ++c) (void)resolve(c);
            dense_mip_texel_lean<false>(src, dst, n, x, y, z, true);  // the 8 children just written
            continue;
        }
        const size_t Dd = (size_t)n + 2, dblock = Dd * Dd * Dd;
        f4 v[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) v[c] = unpack_unorm(resolve(c));
#pragma unroll
        for (int f = 0; f < kFaces; ++f) {
            const size_t di = (size_t)f * dblock + ((size_t)(z + 1) * Dd + (size_t)(y + 1)) * Dd + (size_t)(x + 1);
            ((uint32_t*)dst)[di] = pack_unorm(dense_mip_face(v, f));
        }
    }
}

__device__ void trace_task(const SceneDev& S, uint32_t tile, uint32_t schedule, uint32_t& steps)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int px = (int)(tile % (uint32_t)S.tiles_x) * kTile + (wave & 1) * 8 + (lane & 7);
    const int py = (int)(tile / (uint32_t)S.tiles_x) * kTile + (wave >> 1) * 8 + (lane >> 3);
    if (px >= S.width || py >= S.height) return;
    float4 o;
    if (schedule)
        o = dense::dense_pixel<false, true>(S.data, S.pyr, S.march, px, py, S.width, S.height, S.g_diffuse,
                                            S.g_normal, S.g_depth, steps);
    else
        o = dense::dense_pixel<false, false>(S.data, S.pyr, S.march, px, py, S.width, S.height, S.g_diffuse,
                                             S.g_normal, S.g_depth, steps);
    S.out[(size_t)py * (size_t)S.width + (size_t)px] = o;
}

// The persistent launch: the mip bricks and pixel tiles of every scene, from
// task A.first on (the vertex and raster phases ran before it as flat
// launches). A task waits on its scene's previous phase: mip level m >= 2 on
// level m - 1, a pixel tile on the last level; level 1 reads what the raster
// launch left. Every task a wait depends on sits earlier in the queue, so it
// was taken by a workgroup that is running (a workgroup takes a task only
// while resident, and its own tasks wait on nothing later): waits end.
template <bool kAhead>
__global__ __launch_bounds__(256) void k_batch(Args A)
{
    __shared__ uint32_t s_task;
    __shared__ uint32_t s_go;
    uint32_t steps = 0;
    for (;;) {
        if (threadIdx.x == 0) s_task = A.first + atomicAdd(A.head, 1u);
        __syncthreads();
        const uint32_t ti = s_task;
        if (ti >= A.ntasks) break;
        const Task T = A.tasks[ti];
        const uint32_t type = T.type_scene >> 24, scene = T.type_scene & 0xffffffu;
        const SceneDev& S = A.scenes[scene];
        int dep = -1;
        if (type == kMip && T.level > 1) dep = (int)T.level;  // phase 1 + (level - 1)
        else if (type == kTrace) dep = S.pyr.nlevels;          // the last level, phase 1 + (nlevels - 1)
        if (threadIdx.x == 0) {
            bool ok = !__hip_atomic_load(A.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (ok && dep >= 0)
                ok = wait_count(&A.done[scene * kPhases + dep], S.need[dep], A.abort_flag, A.status, A.wait_ticks);
            s_go = ok ? 1u : 0u;
        }
        __syncthreads();
        if (!s_go) break;
        // Thread 0's acquire (its wait) and the barrier order this workgroup
        // after the producers; a mip brick's writes are ordered before its
        // completion count by the barrier and thread 0's release. No fence
        // beyond those: on this chip an agent-scope fence writes back and
        // invalidates the XCD's L2, which the pixel tiles depend on.
        if (type == kMip) {
            mip_task<true>(S, T.a, (int)T.level, A.mip_brick);
            __syncthreads();
            if (threadIdx.x == 0)
                __hip_atomic_fetch_add(&A.done[scene * kPhases + 1 + (int)T.level], 1u, __ATOMIC_RELEASE,
                                       __HIP_MEMORY_SCOPE_AGENT);
        } else {
            trace_task(S, T.a, kAhead ? 1u : 0u, steps);
            __syncthreads();  // s_task / s_go are rewritten by the next trip
        }
    }
    if (A.samples) {
        unsigned long long v = steps;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if ((threadIdx.x & 63) == 0 && v) atomicAdd(A.samples, v);
    }
}

// Scene of global soup triangle g: the last s with tri_start[s] <= g.
__device__ __forceinline__ uint32_t scene_of(const uint32_t* __restrict__ tri_start, uint32_t nscenes, uint32_t g)
{
    uint32_t lo = 0, hi = nscenes;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tri_start[mid] <= g) lo = mid; else hi = mid;
    }
    return lo;
}

// Raster samples of every triangle of the batch (k_dense_count per scene).
__global__ __launch_bounds__(256) void k_phase_count(const SceneDev* __restrict__ scenes,
                                                     const uint32_t* __restrict__ tri_start, uint32_t nscenes,
                                                     uint32_t ntris, unsigned long long* __restrict__ counts)
{
#pragma clang fp contract(off)
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ntris) return;
    const uint32_t sc = scene_of(tri_start, nscenes, g);
    const SceneDev& S = scenes[sc];
    const uint32_t t = g - tri_start[sc];
    float p[9];
    for (int k = 0; k < 9; ++k) p[k] = S.soup_pos[9u * t + k] - S.grid_min[k % 3];
    const int lo[3] = {0, 0, 0}, hi[3] = {S.res - 1, S.res - 1, S.res - 1};
    TriSetup st;
    unsigned long long n = 0;
    if (tri_setup(p, S.vs, lo, hi, st)) n = (unsigned long long)st.nu * st.nv * (unsigned long long)S.samples;
    counts[g] = n;
}

__global__ void k_phase_total(const unsigned long long* __restrict__ counts,
                              const unsigned long long* __restrict__ offsets, uint32_t n,
                              unsigned long long* __restrict__ total)
{
    total[0] = n ? offsets[n - 1] + counts[n - 1] : 0ull;
}

// Every covered sample of the batch proposes its (scene-local) triangle for
// its voxel; the smallest index wins (k_dense_raster per scene).
__global__ __launch_bounds__(256) void k_phase_raster_flat(const SceneDev* __restrict__ scenes,
                                                           const uint32_t* __restrict__ tri_start, uint32_t nscenes,
                                                           uint32_t ntris, const unsigned long long* __restrict__ offsets,
                                                           const unsigned long long* __restrict__ total)
{
#pragma clang fp contract(off)
    const unsigned long long n = total[0];
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long base = (unsigned long long)blockIdx.x * blockDim.x; base < n; base += stride) {
        const TriRange tr = block_triangles(offsets, ntris, base, min(base + blockDim.x - 1, n - 1));
        // the scenes of the block's first and last triangles bound every lane's scene
        const uint32_t sc_lo = scene_of(tri_start, nscenes, tr.lo), sc_hi = scene_of(tri_start, nscenes, tr.hi);
        const unsigned long long wi = base + threadIdx.x;
        if (wi >= n) continue;
        const uint32_t g = work_triangle(offsets, tr.lo, tr.hi, wi);
        uint32_t sc = sc_lo;
        for (uint32_t hi = sc_hi + 1; hi - sc > 1;) {  // last scene in [sc_lo, sc_hi] starting at or before g
            const uint32_t mid = (sc + hi) >> 1;
            if (tri_start[mid] <= g) sc = mid; else hi = mid;
        }
        const SceneDev& S = scenes[sc];
        const uint32_t t = g - tri_start[sc];
        float p[9];
        for (int k = 0; k < 9; ++k) p[k] = S.soup_pos[9u * t + k] - S.grid_min[k % 3];
        const int rlo[3] = {0, 0, 0}, rhi[3] = {S.res - 1, S.res - 1, S.res - 1};
        TriSetup st;
        if (!tri_setup(p, S.vs, rlo, rhi, st)) continue;
        float fp[3];
        double lam[3];
        if (!tri_sample(p, st, S.vs, S.samples, (uint32_t)(wi - offsets[g]), fp, lam)) continue;
        int ic[3];
        bool inside = true;
        for (int k = 0; k < 3; ++k) {
            const float f = floorf(fp[k] / S.vs);
            if (!(f >= 0.0f) || f >= (float)S.res) inside = false;
            ic[k] = (int)f;
        }
        if (!inside) continue;
        atomicMin(&S.winner[((size_t)ic[2] * S.res + ic[1]) * S.res + ic[0]], t);
    }
}

// The phased schedule: block b runs task first + b (tiles in XCD order).
__global__ __launch_bounds__(256) void k_phase_vertex(const Task* __restrict__ tasks, uint32_t first,
                                                      const SceneDev* __restrict__ scenes)
{
    const Task T = tasks[first + blockIdx.x];
    vertex_task(scenes[T.type_scene & 0xffffffu], T.a, T.b);
}

__global__ __launch_bounds__(256) void k_phase_mip(const Task* __restrict__ tasks, uint32_t first,
                                                   const SceneDev* __restrict__ scenes, uint32_t brick)
{
    const Task T = tasks[first + blockIdx.x];
    mip_task<false>(scenes[T.type_scene & 0xffffffu], T.a, (int)T.level, brick);
}

template <bool kAhead>
__global__ __launch_bounds__(256) void k_phase_trace(const Task* __restrict__ tasks, uint32_t first,
                                                     const SceneDev* __restrict__ scenes, unsigned long long* samples)
{
    const Task T = tasks[first + xcd_remap(blockIdx.x, gridDim.x)];
    uint32_t steps = 0;
    trace_task(scenes[T.type_scene & 0xffffffu], T.a, kAhead ? 1u : 0u, steps);
    if (samples) {
        unsigned long long v = steps;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if ((threadIdx.x & 63) == 0 && v) atomicAdd(samples, v);
    }
}

}  // namespace batch
}  // namespace vct

using namespace vct;
using namespace vct::batch;

struct vct_batch {
    std::vector<vct_batch_scene> scenes;
    std::vector<SceneDev> dev;  // host copy of the scene table
    vct_dense_trace_desc trace{};
    vct_batch_tuning tuning{256, 8, 0, 0};
    int cus = 0;
    DeviceBuffer scene_table, prims, tasks, counters, soup_pos, soup_mat, data, winner;
    DeviceBuffer tri_start, counts, offsets, total, scan_tmp;  // the phased raster
    uint32_t ntris = 0;
    uint32_t ntasks = 0;
    uint32_t ndone = 0;
    struct Phase {
        uint32_t type, level, first, count;
    };
    std::vector<Phase> phases;  // contiguous runs of the task table, in order
    long long wait_ticks = 0;
    std::vector<size_t> data_off, winner_off;
    hipStream_t capture = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    unsigned long long* graph_samples = nullptr;  // samples pointer the graph was captured with
    bool graph_valid = false;
};

static void drop_graph(vct_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_valid = false;
}

// The shape of one scene as the planner sees it.
struct PlanScene {
    uint32_t tri_count, res, width, height;
};

// Pyramid levels of a res^3 grid: log2(res) + 1.
static int plan_levels(uint32_t res)
{
    int nl = 0;
    while ((1u << nl) < res) ++nl;
    return nl + 1;
}

// Task table for raster chunk rc and mip brick B, and the task count of every
// (scene, phase). Phases in queue order: vertex chunks of every scene, raster
// chunks, mip level 1 of every scene, level 2, ..., pixel tiles. Within a
// phase, larger work first for voxelization and mips; trace tiles of the
// scenes whose mips finish first (smaller grids) first. Pure host code.
static void plan_tasks(const PlanScene* sc, uint32_t ns, uint32_t rc, uint32_t B, std::vector<Task>& t,
                       std::vector<uint32_t>& need)
{
    std::vector<uint32_t> by_tris(ns), by_res_desc(ns), by_res_asc(ns);
    for (uint32_t i = 0; i < ns; ++i) by_tris[i] = by_res_desc[i] = by_res_asc[i] = i;
    std::stable_sort(by_tris.begin(), by_tris.end(),
                     [&](uint32_t x, uint32_t y) { return sc[x].tri_count > sc[y].tri_count; });
    std::stable_sort(by_res_desc.begin(), by_res_desc.end(), [&](uint32_t x, uint32_t y) { return sc[x].res > sc[y].res; });
    std::stable_sort(by_res_asc.begin(), by_res_asc.end(), [&](uint32_t x, uint32_t y) {
        return sc[x].res != sc[y].res ? sc[x].res < sc[y].res : sc[x].tri_count < sc[y].tri_count;
    });
    t.clear();
    need.assign((size_t)ns * kPhases, 0u);
    for (uint32_t s : by_tris)
        for (uint32_t a = 0; a < sc[s].tri_count; a += 256, ++need[(size_t)s * kPhases + 0])
            t.push_back({(kVertex << 24) | s, a, std::min(a + 256u, sc[s].tri_count), 0});
    for (uint32_t s : by_tris)
        for (uint32_t a = 0; a < sc[s].tri_count; a += rc, ++need[(size_t)s * kPhases + 1])
            t.push_back({(kRaster << 24) | s, a, std::min(a + rc, sc[s].tri_count), 0});
    int max_levels = 0;
    for (uint32_t s = 0; s < ns; ++s) max_levels = std::max(max_levels, plan_levels(sc[s].res));
    for (int m = 1; m < max_levels; ++m)
        for (uint32_t s : by_res_desc) {
            if (m >= plan_levels(sc[s].res)) continue;
            const uint32_t n = sc[s].res >> m, be = std::min(n, B), nb = (n + be - 1) / be;
            for (uint32_t k = 0; k < nb * nb * nb; ++k, ++need[(size_t)s * kPhases + 1 + m])
                t.push_back({(kMip << 24) | s, k, 0, (uint32_t)m});
        }
    for (uint32_t s : by_res_asc) {
        const uint32_t tiles = ((sc[s].width + kTile - 1) / kTile) * ((sc[s].height + kTile - 1) / kTile);
        for (uint32_t k = 0; k < tiles; ++k) t.push_back({(kTrace << 24) | s, k, 0, 0});
    }
}

// Task table for the current tuning, uploaded with the scene table (whose
// need[] counts it sets).
static vct_status build_tasks(vct_batch* b)
{
    const uint32_t ns = (uint32_t)b->scenes.size();
    std::vector<PlanScene> shapes(ns);
    for (uint32_t i = 0; i < ns; ++i)
        shapes[i] = {b->dev[i].tri_count, (uint32_t)b->dev[i].res, (uint32_t)b->dev[i].width,
                     (uint32_t)b->dev[i].height};
    std::vector<Task> t;
    std::vector<uint32_t> need;
    plan_tasks(shapes.data(), ns, b->tuning.raster_chunk, b->tuning.mip_brick, t, need);
    for (uint32_t i = 0; i < ns; ++i)
        for (int p = 0; p < kPhases; ++p) b->dev[i].need[p] = need[(size_t)i * kPhases + p];
    b->ntasks = (uint32_t)t.size();
    b->phases.clear();
    for (uint32_t i = 0; i < b->ntasks; ++i) {
        const uint32_t type = t[i].type_scene >> 24, level = type == kMip ? t[i].level : 0;
        if (b->phases.empty() || b->phases.back().type != type || b->phases.back().level != level)
            b->phases.push_back({type, level, i, 0});
        ++b->phases.back().count;
    }
    VCT_TRY(ensure(b->tasks, sizeof(Task) * t.size()));
    VCT_HIP(hipMemcpy(b->tasks.ptr, t.data(), sizeof(Task) * t.size(), hipMemcpyHostToDevice));
    VCT_HIP(hipMemcpy(b->scene_table.ptr, b->dev.data(), sizeof(SceneDev) * ns, hipMemcpyHostToDevice));
    return VCT_OK;
}

static vct_status capture_graph(vct_batch* b, unsigned long long* samples)
{
    drop_graph(b);
    uint32_t* ctr = (uint32_t*)b->counters.ptr;
    Args A;
    A.tasks = (const Task*)b->tasks.ptr;
    A.ntasks = b->ntasks;
    A.scenes = (const SceneDev*)b->scene_table.ptr;
    A.head = ctr;
    A.abort_flag = ctr + 1;
    A.status = ctr + 2;
    A.done = ctr + 4;
    A.samples = samples;
    A.mip_brick = b->tuning.mip_brick;
    A.schedule = b->tuning.trace_schedule;
    A.wait_ticks = b->wait_ticks;
    const unsigned grid = (unsigned)(b->cus * (int)b->tuning.workgroups_per_cu);
    A.first = b->ntasks;
    for (const vct_batch::Phase& p : b->phases)
        if (p.type == kMip || p.type == kTrace) {
            A.first = p.first;
            break;
        }
    VCT_HIP(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(k_batch_reset, dim3((b->ndone + 255) / 256), dim3(256), 0, b->capture, A.head, A.done, b->ndone,
                       A.abort_flag);
    // Both schedules: the vertex chunks, then the raster of all scenes'
    // triangles as one sample-parallel pass over the batch; then one flat
    // launch per mip level and one for the pixel tiles, or (grid > 0) one
    // persistent launch pulling the mip bricks and pixel tiles from the queue.
    for (const vct_batch::Phase& p : b->phases) {
        const dim3 g(p.count);
        if (p.type == kVertex) {
            hipLaunchKernelGGL(k_phase_vertex, g, dim3(256), 0, b->capture, A.tasks, p.first, A.scenes);
        } else if (p.type == kRaster) {
            const uint32_t ns = (uint32_t)b->scenes.size(), nt = b->ntris;
            auto* counts = (unsigned long long*)b->counts.ptr;
            auto* offsets = (unsigned long long*)b->offsets.ptr;
            auto* total = (unsigned long long*)b->total.ptr;
            hipLaunchKernelGGL(k_phase_count, dim3((nt + 255) / 256), dim3(256), 0, b->capture, A.scenes,
                               (const uint32_t*)b->tri_start.ptr, ns, nt, counts);
            size_t tmp = b->scan_tmp.bytes;
            hipError_t e = hipcub::DeviceScan::ExclusiveSum(b->scan_tmp.ptr, tmp, counts, offsets, (int)nt, b->capture);
            if (e != hipSuccess) {
                hipGraph_t g0 = nullptr;
                (void)hipStreamEndCapture(b->capture, &g0);
                if (g0) (void)hipGraphDestroy(g0);
                return fail(VCT_ERR_HIP, "batched raster scan: %s", hipGetErrorString(e));
            }
            hipLaunchKernelGGL(k_phase_total, dim3(1), dim3(1), 0, b->capture, counts, offsets, nt, total);
            hipLaunchKernelGGL(k_phase_raster_flat, dim3((unsigned)(b->cus * 32)), dim3(256), 0, b->capture, A.scenes,
                               (const uint32_t*)b->tri_start.ptr, ns, nt, offsets, total);
        } else if (grid) {
            break;  // the rest is the persistent launch below
        } else if (p.type == kMip) {
            hipLaunchKernelGGL(k_phase_mip, g, dim3(256), 0, b->capture, A.tasks, p.first, A.scenes, A.mip_brick);
        } else if (A.schedule) {
            hipLaunchKernelGGL(k_phase_trace<true>, g, dim3(256), 0, b->capture, A.tasks, p.first, A.scenes, samples);
        } else {
            hipLaunchKernelGGL(k_phase_trace<false>, g, dim3(256), 0, b->capture, A.tasks, p.first, A.scenes, samples);
        }
    }
    if (grid && A.schedule) hipLaunchKernelGGL(k_batch<true>, dim3(grid), dim3(256), 0, b->capture, A);
    else if (grid) hipLaunchKernelGGL(k_batch<false>, dim3(grid), dim3(256), 0, b->capture, A);
    hipError_t e = hipStreamEndCapture(b->capture, &b->graph);
    if (e != hipSuccess) {
        b->graph = nullptr;
        return fail(VCT_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    }
    VCT_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_samples = samples;
    b->graph_valid = true;
    return VCT_OK;
}

extern "C" {

static vct_status batch_init(vct_batch* b, const vct_batch_scene* scenes, uint32_t count,
                             const vct_dense_trace_desc* trace);

vct_status vct_batch_create(vct_batch** out, const vct_batch_scene* scenes, uint32_t count,
                            const vct_dense_trace_desc* trace)
{
    if (!out || !scenes || !trace || !count) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument or empty batch");
    *out = nullptr;
    if (count > 0xffffffu) return fail(VCT_ERR_INVALID_ARGUMENT, "too many scenes");
    vct_batch* b = new (std::nothrow) vct_batch();
    if (!b) return fail(VCT_ERR_OUT_OF_MEMORY, "host allocation failed");
    const vct_status s = batch_init(b, scenes, count, trace);
    if (s != VCT_OK) {
        vct_batch_destroy(b);
        return s;
    }
    *out = b;
    return VCT_OK;
}

}  // extern "C"

// Validates the scenes, lays out every scene's soup, pyramid and winners in
// shared allocations, and builds the task table for the default tuning.
static vct_status batch_init(vct_batch* b, const vct_batch_scene* scenes, uint32_t count,
                             const vct_dense_trace_desc* trace)
{
    b->scenes.assign(scenes, scenes + count);
    b->trace = *trace;
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
        return fail(VCT_ERR_HIP, "device query failed");
    b->wait_ticks = (long long)(kWaitSeconds * 1e3 * (double)khz);
    // per-scene validation and sizes
    std::vector<Prim> prims;
    size_t tris_total = 0, data_total = 0, winner_total = 0;
    b->dev.resize(count);
    for (uint32_t i = 0; i < count; ++i) {
        const vct_batch_scene& sc = scenes[i];
        const vct_mesh& m = sc.mesh;
        if (!m.positions || !m.indices || !m.node_matrices || !m.materials || (!m.primitives && m.primitive_count))
            return fail(VCT_ERR_INVALID_ARGUMENT, "scene %u: mesh buffers must be non-null", i);
        const uint32_t r = sc.grid.resolution;
        if (r < 2 || r > 1024 || (r & (r - 1)))
            return fail(VCT_ERR_INVALID_ARGUMENT, "scene %u: resolution must be a power of two in [2, 1024]", i);
        if (sc.grid.format != VCT_FORMAT_RGBA8) return fail(VCT_ERR_UNSUPPORTED, "scene %u: batches are RGBA8", i);
        if (sc.grid.raster_samples != 1 && sc.grid.raster_samples != 4)
            return fail(VCT_ERR_INVALID_ARGUMENT, "scene %u: raster_samples", i);
        if (!(sc.grid.extent > 0.0f)) return fail(VCT_ERR_INVALID_ARGUMENT, "scene %u: extent", i);
        if (!sc.gbuffer.diffuse || !sc.gbuffer.normal || !sc.gbuffer.depth || !sc.gbuffer.width || !sc.gbuffer.height ||
            !sc.out)
            return fail(VCT_ERR_INVALID_ARGUMENT, "scene %u: G-buffer and output are required", i);
        {
            std::vector<int32_t> tex(3u * (m.material_count ? m.material_count : 1u));
            VCT_TRY(read_texture_indices(nullptr, m.materials, m.material_count, tex.data()));
            for (uint32_t k = 0; k < m.primitive_count; ++k)
                if (m.primitives[k].index_count && m.primitives[k].material_index < m.material_count &&
                    (tex[3u * m.primitives[k].material_index] >= 0 ||
                     tex[3u * m.primitives[k].material_index + 1] >= 0 ||
                     tex[3u * m.primitives[k].material_index + 2] >= 0))
                    return fail(VCT_ERR_UNSUPPORTED,
                                "scene %u primitive %u: the material references a base-colour, emissive or "
                                "occlusion texture; the voxel passes use material factors only",
                                i, k);
        }
        SceneDev& d = b->dev[i];
        std::memset(&d, 0, sizeof(d));
        d.positions = m.positions;
        d.indices = m.indices;
        d.nodes = m.node_matrices;
        d.materials = m.materials;
        const uint32_t pbase = (uint32_t)prims.size();
        uint64_t nt = 0;
        for (uint32_t k = 0; k < m.primitive_count; ++k) {
            const vct_primitive& p = m.primitives[k];
            if (p.index_count % 3u || (uint64_t)p.first_index + p.index_count > m.index_count ||
                p.instance_index >= m.instance_count || p.material_index >= m.material_count)
                return fail(VCT_ERR_INVALID_ARGUMENT, "scene %u primitive %u out of range", i, k);
            if (!p.index_count) continue;
            prims.push_back({p.first_index, p.index_count / 3u, p.vertex_offset, p.instance_index, p.material_index,
                             (uint32_t)nt});
            nt += p.index_count / 3u;
        }
        if (nt > 0xffffffffull / 9u) return fail(VCT_ERR_CAPACITY, "scene %u: too many triangles", i);
        d.nprims = (uint32_t)prims.size() - pbase;
        d.prims = (const Prim*)(uintptr_t)pbase;  // rebased below
        d.tri_count = (uint32_t)nt;
        d.soup_pos = (float*)(uintptr_t)tris_total;
        tris_total += nt;
        d.res = (int)r;
        d.samples = (int)sc.grid.raster_samples;
        d.vs = sc.grid.extent / (float)r;
        for (int k = 0; k < 3; ++k) d.grid_min[k] = sc.grid.grid_min[k];
        int nl = 0;
        while ((1u << nl) < r) ++nl;
        d.pyr.nlevels = nl + 1;
        if (d.pyr.nlevels > dense::kMaxLevels) return fail(VCT_ERR_UNSUPPORTED, "scene %u: too many levels", i);
        size_t tex = 0;
        for (int lv = 0; lv < d.pyr.nlevels; ++lv) {
            const int n = (int)r >> lv;
            d.pyr.offset[lv] = tex;
            d.pyr.block[lv] = (unsigned long long)(n + 2) * (n + 2) * (n + 2);
            d.pyr.n[lv] = n;
            d.pyr.inv_vs[lv] = (float)n / sc.grid.extent;
            tex += d.pyr.block[lv] * (lv == 0 ? 1u : 6u);
        }
        d.pyr.vs = d.vs;
        b->data_off.push_back(data_total);
        data_total += tex;
        b->winner_off.push_back(winner_total);
        winner_total += (size_t)r * r * r;
        vct_dense_trace_desc td = *trace;
        vct_dense tmp;  // only the fields make_params reads
        tmp.desc = sc.grid;
        tmp.res = (int)r;
        tmp.nlevels = d.pyr.nlevels;
        for (int lv = 0; lv < d.pyr.nlevels; ++lv) {
            tmp.offset[lv] = d.pyr.offset[lv];
            tmp.block[lv] = d.pyr.block[lv];
        }
        VCT_TRY(dense::make_params(&tmp, &sc.camera, &td, &d.pyr, &d.march));
        d.g_diffuse = sc.gbuffer.diffuse;
        d.g_normal = sc.gbuffer.normal;
        d.g_depth = sc.gbuffer.depth;
        d.width = (int)sc.gbuffer.width;
        d.height = (int)sc.gbuffer.height;
        d.tiles_x = (d.width + kTile - 1) / kTile;
        d.out = (float4*)sc.out;
    }
    VCT_TRY(ensure(b->prims, sizeof(Prim) * (prims.size() ? prims.size() : 1)));
    VCT_TRY(ensure(b->soup_pos, 36u * (tris_total ? tris_total : 1)));
    VCT_TRY(ensure(b->soup_mat, 4u * (tris_total ? tris_total : 1)));
    VCT_TRY(ensure(b->data, 4u * data_total));
    VCT_TRY(ensure(b->winner, 4u * winner_total));
    b->ndone = count * kPhases;
    VCT_TRY(ensure(b->counters, 4u * (4 + (size_t)b->ndone)));
    VCT_TRY(ensure(b->scene_table, sizeof(SceneDev) * count));
    if (!prims.empty())
        VCT_HIP(hipMemcpy(b->prims.ptr, prims.data(), sizeof(Prim) * prims.size(), hipMemcpyHostToDevice));
    if (tris_total > 0x7fffffffull) return fail(VCT_ERR_CAPACITY, "more than 2^31 triangles in one batch");
    std::vector<uint32_t> tri_start(count);
    for (uint32_t i = 0; i < count; ++i) {
        SceneDev& d = b->dev[i];
        d.prims = (const Prim*)b->prims.ptr + (uintptr_t)d.prims;
        const size_t tb = (size_t)(uintptr_t)d.soup_pos;
        tri_start[i] = (uint32_t)tb;
        d.soup_pos = (float*)b->soup_pos.ptr + 9u * tb;
        d.soup_mat = (uint32_t*)b->soup_mat.ptr + tb;
        d.data = (uint32_t*)b->data.ptr + b->data_off[i];
        d.winner = (uint32_t*)b->winner.ptr + b->winner_off[i];
    }
    // the phased raster's tables
    b->ntris = (uint32_t)tris_total;
    VCT_TRY(ensure(b->tri_start, 4u * count));
    VCT_HIP(hipMemcpy(b->tri_start.ptr, tri_start.data(), 4u * count, hipMemcpyHostToDevice));
    VCT_TRY(ensure(b->counts, 8u * (tris_total ? tris_total : 1)));
    VCT_TRY(ensure(b->offsets, 8u * (tris_total ? tris_total : 1)));
    VCT_TRY(ensure(b->total, 16));
    size_t scan_bytes = 0;
    VCT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (unsigned long long*)b->counts.ptr,
                                             (unsigned long long*)b->offsets.ptr, (int)tris_total, (hipStream_t)0));
    VCT_TRY(ensure(b->scan_tmp, scan_bytes ? scan_bytes : 16));
    // borders stay zero forever; winners start empty and every launch leaves them empty
    VCT_HIP(hipMemset(b->data.ptr, 0, 4u * data_total));
    VCT_HIP(hipMemset(b->winner.ptr, 0xff, 4u * winner_total));
    VCT_HIP(hipMemset(b->counters.ptr, 0, 4u * (4 + (size_t)b->ndone)));
    VCT_HIP(hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking));
    VCT_TRY(build_tasks(b));
    VCT_HIP(hipDeviceSynchronize());
    return VCT_OK;
}

extern "C" {

vct_status vct_batch_destroy(vct_batch* b)
{
    if (!b) return VCT_OK;
    (void)hipDeviceSynchronize();
    drop_graph(b);
    if (b->capture) hipStreamDestroy(b->capture);
    DeviceBuffer* bufs[] = {&b->scene_table, &b->prims,  &b->tasks,  &b->counters, &b->soup_pos, &b->soup_mat,
                            &b->data,        &b->winner, &b->tri_start, &b->counts, &b->offsets, &b->total,
                            &b->scan_tmp};
    for (DeviceBuffer* p : bufs) release(*p);
    delete b;
    return VCT_OK;
}

vct_status vct_batch_set_tuning(vct_batch* b, const vct_batch_tuning* t)
{
    if (!b || !t) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    if (t->raster_chunk != 64 && t->raster_chunk != 128 && t->raster_chunk != 256)
        return fail(VCT_ERR_INVALID_ARGUMENT, "raster_chunk must be 64, 128 or 256");
    if (t->mip_brick != 4 && t->mip_brick != 8) return fail(VCT_ERR_INVALID_ARGUMENT, "mip_brick must be 4 or 8");
    if (t->trace_schedule > 1) return fail(VCT_ERR_INVALID_ARGUMENT, "trace_schedule must be 0 or 1");
    if (t->workgroups_per_cu > 4) return fail(VCT_ERR_INVALID_ARGUMENT, "workgroups_per_cu must be 0..4");
    VCT_HIP(hipDeviceSynchronize());
    const bool retask = t->raster_chunk != b->tuning.raster_chunk || t->mip_brick != b->tuning.mip_brick;
    b->tuning = *t;
    drop_graph(b);
    if (retask) VCT_TRY(build_tasks(b));
    return VCT_OK;
}

vct_status vct_batch_get_tuning(const vct_batch* b, vct_batch_tuning* t)
{
    if (!b || !t) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    *t = b->tuning;
    return VCT_OK;
}

vct_status vct_batch_run(vct_batch* b, vct_stream stream, unsigned long long* samples_out)
{
    if (!b) return fail(VCT_ERR_INVALID_ARGUMENT, "null batch");
    if (!b->graph_valid || b->graph_samples != samples_out) VCT_TRY(capture_graph(b, samples_out));
    VCT_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
    return VCT_OK;
}

vct_status vct_batch_status(vct_batch* b)
{
    if (!b) return fail(VCT_ERR_INVALID_ARGUMENT, "null batch");
    VCT_HIP(hipDeviceSynchronize());
    uint32_t st = 0;
    VCT_HIP(hipMemcpy(&st, (uint32_t*)b->counters.ptr + 2, 4, hipMemcpyDeviceToHost));
    if (st) {
        const uint32_t zero = 0;
        VCT_HIP(hipMemcpy((uint32_t*)b->counters.ptr + 2, &zero, 4, hipMemcpyHostToDevice));
        // an aborted launch may have skipped the mip level-1 tasks that return
        // the winners to empty; the next launch needs them empty
        size_t winner_total = 0;
        for (const SceneDev& d : b->dev) winner_total += (size_t)d.res * d.res * d.res;
        VCT_HIP(hipMemset(b->winner.ptr, 0xff, 4u * winner_total));
        VCT_HIP(hipDeviceSynchronize());
        return fail(VCT_ERR_HIP, "a batched task gave up waiting on its dependency; results are incomplete");
    }
    return VCT_OK;
}

vct_status vct_batch_plan(const vct_batch_scene_shape* shapes, uint32_t count, const vct_batch_tuning* tuning,
                          uint32_t* tasks, size_t cap_tasks, size_t* n_tasks, uint32_t* need)
{
    if (!shapes || !count || !tuning || !n_tasks) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    if (tuning->raster_chunk != 64 && tuning->raster_chunk != 128 && tuning->raster_chunk != 256)
        return fail(VCT_ERR_INVALID_ARGUMENT, "raster_chunk must be 64, 128 or 256");
    if (tuning->mip_brick != 4 && tuning->mip_brick != 8) return fail(VCT_ERR_INVALID_ARGUMENT, "mip_brick must be 4 or 8");
    std::vector<PlanScene> sc(count);
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t r = shapes[i].resolution;
        if (r < 2 || r > 1024 || (r & (r - 1)))
            return fail(VCT_ERR_INVALID_ARGUMENT, "scene %u: resolution must be a power of two in [2, 1024]", i);
        if (!shapes[i].width || !shapes[i].height) return fail(VCT_ERR_INVALID_ARGUMENT, "scene %u: image size", i);
        sc[i] = {shapes[i].triangles, r, shapes[i].width, shapes[i].height};
    }
    std::vector<Task> t;
    std::vector<uint32_t> nd;
    plan_tasks(sc.data(), count, tuning->raster_chunk, tuning->mip_brick, t, nd);
    *n_tasks = t.size();
    if (need) std::memcpy(need, nd.data(), sizeof(uint32_t) * nd.size());
    if (tasks) {
        if (cap_tasks < t.size()) return fail(VCT_ERR_CAPACITY, "the plan holds %zu tasks", t.size());
        std::memcpy(tasks, t.data(), sizeof(Task) * t.size());
    }
    return VCT_OK;
}

vct_status vct_batch_autotune(vct_batch* b, vct_stream stream, uint32_t reps, vct_batch_tuning* best,
                              float* best_ms)
{
    if (!b) return fail(VCT_ERR_INVALID_ARGUMENT, "null batch");
    if (reps == 0) reps = 3;
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    VCT_HIP(hipEventCreate(&e0));
    VCT_HIP(hipEventCreate(&e1));
    auto time_it = [&](const vct_batch_tuning& t, float* ms) -> vct_status {
        VCT_TRY(vct_batch_set_tuning(b, &t));
        VCT_TRY(vct_batch_run(b, stream, nullptr));  // capture + warm-up
        std::vector<float> v;
        for (uint32_t r = 0; r < reps; ++r) {
            VCT_HIP(hipEventRecord(e0, s));
            VCT_TRY(vct_batch_run(b, stream, nullptr));
            VCT_HIP(hipEventRecord(e1, s));
            VCT_HIP(hipEventSynchronize(e1));
            float x = 0.0f;
            VCT_HIP(hipEventElapsedTime(&x, e0, e1));
            v.push_back(x);
        }
        std::sort(v.begin(), v.end());
        *ms = v[v.size() / 2];
        return vct_batch_status(b);
    };
    // coordinate descent over the knobs, two sweeps (the raster chunk no
    // longer changes the launches: both schedules raster all scenes' samples
    // in one sample-parallel pass)
    static const uint32_t bricks[] = {4, 8}, scheds[] = {0, 1}, occ[] = {0, 1, 2, 4};
    vct_batch_tuning cur = b->tuning;
    float cur_ms = 0.0f;
    vct_status st = time_it(cur, &cur_ms);
    for (int sweep = 0; st == VCT_OK && sweep < 2; ++sweep) {
        for (int knob = 0; st == VCT_OK && knob < 3; ++knob) {
            const uint32_t* vals = knob == 0 ? bricks : knob == 1 ? scheds : occ;
            const int nv = knob == 0 ? 2 : knob == 1 ? 2 : 4;
            for (int k = 0; st == VCT_OK && k < nv; ++k) {
                vct_batch_tuning t = cur;
                uint32_t* field = knob == 0 ? &t.mip_brick : knob == 1 ? &t.trace_schedule : &t.workgroups_per_cu;
                if (*field == vals[k]) continue;
                *field = vals[k];
                float ms = 0.0f;
                st = time_it(t, &ms);
                if (st == VCT_OK && ms < cur_ms) {
                    cur = t;
                    cur_ms = ms;
                }
            }
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (st != VCT_OK) return st;
    VCT_TRY(vct_batch_set_tuning(b, &cur));
    if (best) *best = cur;
    if (best_ms) *best_ms = cur_ms;
    return VCT_OK;
}

vct_status vct_batch_level(vct_batch* b, uint32_t scene, uint32_t level, void** ptr, size_t* bytes, uint32_t* faces)
{
    if (!b || !ptr) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    if (scene >= b->dev.size()) return fail(VCT_ERR_INVALID_ARGUMENT, "scene out of range");
    const SceneDev& d = b->dev[scene];
    if (level >= (uint32_t)d.pyr.nlevels) return fail(VCT_ERR_INVALID_ARGUMENT, "level out of range");
    *ptr = (uint32_t*)d.data + d.pyr.offset[level];
    if (bytes) *bytes = 4u * d.pyr.block[level] * (level == 0 ? 1u : 6u);
    if (faces) *faces = level == 0 ? 1u : 6u;
    return VCT_OK;
}

vct_status vct_batch_read_level(vct_batch* b, vct_stream stream, uint32_t scene, uint32_t level, void* dst,
                                size_t bytes)
{
    void* p;
    size_t n;
    VCT_TRY(vct_batch_level(b, scene, level, &p, &n, nullptr));
    if (!dst || bytes != n) return fail(VCT_ERR_INVALID_ARGUMENT, "buffer is %zu bytes, level is %zu", bytes, n);
    VCT_HIP(hipMemcpyAsync(dst, p, n, hipMemcpyDeviceToHost, (hipStream_t)stream));
    VCT_HIP(hipStreamSynchronize((hipStream_t)stream));
    return VCT_OK;
}

}  // extern "C"
