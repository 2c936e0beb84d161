"""GPU parity of the sparse voxel octree path (K9-K16) against the oracle.

The fragment list and the node pool must be bit-identical to the oracle's
serialisation (fragments in rasterization order, child blocks in pool-index
order). Independently of both, the octree topology is checked against brute
force from the set of occupied cells: the flagged nodes of each depth are the
distinct ancestors of the fragments, every child pointer is a unique
multiple of 8, and the pool holds 8 + 8 * (flagged interior nodes) entries.
The octree cone trace must match the oracle to 1e-3 relative L2 error.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_api as O

pytestmark = pytest.mark.gpu

LEVEL = 6  # 64^3


def _scene():
    from vk_voxel_cone_tracing_amd import scenes

    return scenes.room(1, scale=1.0, n_boxes=16, n_spheres=2, thin_total=6000, emissive_boxes=1)


def _run(quirks, samples=1, level=LEVEL):
    import torch
    import vk_voxel_cone_tracing_amd as V
    from vk_voxel_cone_tracing_amd import scenes

    scene = _scene()
    pos, nrm, mat = scene.world_triangles()
    ctx = V.Context(resolution=16, levels=2, quirks=quirks, raster_samples=samples, downsample_region_size=1,
                    extent_l0=1.0)
    ds = V.DeviceScene(scene)
    light = scenes.light(1, scale=1.0)
    depth = O.raycast_shadow(pos, nrm, mat, light, 128, 128)
    sm = V.shadow_struct(torch.from_numpy(depth).cuda())
    ld, ls = V.light_structs(light)
    nfrag = C.c_uint32(0)
    V._abi.call("vct_svo_voxelize", ctx._h, None, C.byref(ds.mesh), C.byref(sm), C.byref(ld), C.byref(ls), level,
                C.byref(nfrag))
    n = C.c_size_t(0)
    frags = np.zeros((nfrag.value, 2), np.uint32)
    V._abi.call("vct_svo_read_fragments", ctx._h, None, frags.ctypes.data_as(C.c_void_p), nfrag.value, C.byref(n))
    V._abi.call("vct_svo_build", ctx._h, None)
    V._abi.call("vct_svo_read_nodes", ctx._h, None, None, 0, C.byref(n))
    nodes = np.zeros((n.value, 2), np.uint32)
    V._abi.call("vct_svo_read_nodes", ctx._h, None, nodes.ctypes.data_as(C.c_void_p), n.value, C.byref(n))
    # oracle
    L = O.lib()
    cap = 4 * nfrag.value + 16
    ofr = np.zeros((cap, 2), np.uint32)
    osm = O.Shadow(depth.ctypes.data, 128, 128)
    bbmin = (C.c_float * 3)(*scene.bb_min.tolist())
    bbmax = (C.c_float * 3)(*scene.bb_max.tolist())
    onf = L.orc_svo_fragments(C.byref(O.tris_struct(pos, nrm, mat)), O.materials_struct(scene), C.byref(osm),
                              C.byref(O.light_struct(light)), bbmin, bbmax, 1 << level, samples, quirks,
                              O.vptr(ofr), cap)
    opool = np.zeros((max(1000000, 8 * onf), 2), np.uint32)
    ocount = C.c_uint32(0)
    assert L.orc_svo_build(O.vptr(ofr), onf, level, quirks, O.vptr(opool), len(opool), C.byref(ocount)) == 0
    return dict(ctx=ctx, ds=ds, sm=sm, ld=ld, ls=ls, scene=scene, light=light, depth=depth, frags=frags,
                nodes=nodes, ofrags=ofr[:onf], onodes=opool[:ocount.value], pos=pos, nrm=nrm, mat=mat)


@pytest.mark.parametrize("quirks,samples", [(0x3F, 1), (0x00, 1), (0x3F, 4)])
def test_fragment_list_and_pool_bit_exact(quirks, samples):
    r = _run(quirks, samples)
    print(f"quirks={quirks:#x} samples={samples}: {len(r['frags'])} fragments, {len(r['nodes'])} nodes")
    assert len(r["frags"]) > 10000
    np.testing.assert_array_equal(r["frags"], r["ofrags"])
    np.testing.assert_array_equal(r["nodes"], r["onodes"])


@pytest.mark.parametrize("level,quirks", [(1, 0x3F), (2, 0x00), (9, 0x3F), (9, 0x00)])
def test_fragment_list_and_pool_bit_exact_at_depth_extremes(level, quirks):
    # the shallowest octrees (one and two levels) and a 512^3 one
    r = _run(quirks, 1, level)
    print(f"level={level} quirks={quirks:#x}: {len(r['frags'])} fragments, {len(r['nodes'])} nodes")
    assert len(r["frags"]) > 0 and len(r["nodes"]) >= 8
    np.testing.assert_array_equal(r["frags"], r["ofrags"])
    np.testing.assert_array_equal(r["nodes"], r["onodes"])


def _decode(frags, half):
    x = frags[:, 0] & 0xFFF
    y = (frags[:, 0] >> 12) & 0xFFF
    z = ((frags[:, 0] >> 24) & 0xFF) | ((frags[:, 1] >> 20) & 0xF00)
    p = np.stack([x, y, z], 1).astype(np.int64)
    return p >> 1 if half else p


@pytest.mark.parametrize("quirks", [0x3F, 0x00])
def test_topology_against_distinct_ancestors(quirks):
    r = _run(quirks)
    nodes = r["nodes"]
    half = bool(quirks & 0x08)
    p = _decode(r["frags"], half)
    R = 1 << LEVEL
    # brute force: path digits of each fragment; slot = z | x<<1 | y<<2 per depth
    paths = []
    for d in range(1, LEVEL + 1):
        res = R >> d
        paths.append(None)
    cur = p.copy()
    keys = np.zeros(len(p), np.int64)
    distinct = []
    res = R
    for d in range(1, LEVEL + 1):
        res >>= 1
        c = (cur >= res).astype(np.int64)
        slot = c[:, 2] | (c[:, 0] << 1) | (c[:, 1] << 2)
        cur = cur - c * res
        keys = keys * 8 + slot
        distinct.append(len(np.unique(keys)))
    flagged = (nodes[:, 0] & 0x80000000) != 0
    child = nodes[:, 0] & 0x7FFFFFFF
    # BFS over the pool by depth
    level_nodes = [np.arange(8)]
    for d in range(1, LEVEL + 1):
        ids = level_nodes[-1]
        assert flagged[ids].sum() == distinct[d - 1], (d, flagged[ids].sum(), distinct[d - 1])
        if d < LEVEL:
            kids = child[ids][flagged[ids]]
            assert np.all(kids % 8 == 0) and len(np.unique(kids)) == len(kids)
            level_nodes.append((kids[:, None] + np.arange(8)[None, :]).reshape(-1))
        else:
            assert np.all(child[ids] == 0)  # leaves are flagged but never subdivided
    interior = sum(distinct[:-1])
    assert len(nodes) == 8 + 8 * interior


def test_svo_trace_matches_oracle():
    import torch
    import vk_voxel_cone_tracing_amd as V
    from vk_voxel_cone_tracing_amd import scenes

    r = _run(0x3F)
    scene = r["scene"]
    W, H = 40, 30
    vp, vp_inv, eye = scenes.camera(1, W, H, scale=1.0)
    g = O.raycast_gbuffer(scene, r["pos"], r["nrm"], r["mat"], vp, vp_inv, W, H)
    gd = {k: torch.from_numpy(v.view(np.int32) if v.dtype == np.uint32 else
                              (v.view(np.int16) if v.dtype == np.uint16 else v)).cuda() for k, v in g.items()}
    vs0 = 1.0 / 64.0
    center = [0.5, 0.5, 0.5]
    for mode in (8, 7):
        desc = V._abi.ConeTracingDesc(V._abi.f3(*center), mode, vs0, 64.0, 1.0, 8.0, 0.5, 1.0, 3.0, 2.0, 0)
        dout = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        sout = torch.zeros_like(dout)
        bbmin = (C.c_float * 3)(*scene.bb_min.tolist())
        bbmax = (C.c_float * 3)(*scene.bb_max.tolist())
        V._abi.call("vct_svo_trace", r["ctx"]._h, None, C.byref(V.gbuffer_struct(gd)),
                    C.byref(V.camera_ubo(vp, vp_inv, eye)), C.byref(desc), C.byref(r["sm"]), C.byref(r["ld"]),
                    C.byref(r["ls"]), bbmin, bbmax, C.c_void_p(dout.data_ptr()), C.c_void_p(sout.data_ptr()), None)
        torch.cuda.synchronize()
        fields = dict(volume_center=center, rendering_mode=mode, voxel_size=vs0, volume_dimension=64.0,
                      trace_start_offset=1.0, indirect_diffuse_intensity=8.0, ambient_occlusion_factor=0.5,
                      min_trace_step_factor=1.0, indirect_specular_intensity=3.0, occlusion_decay=2.0,
                      enable_32_cones=0)
        p = O.trace_params(fields, vp_inv, eye, 6)
        rd = np.zeros((H, W, 4), np.float32)
        rs = np.zeros_like(rd)
        osm = O.Shadow(r["depth"].ctypes.data, 128, 128)
        pool = np.ascontiguousarray(r["onodes"])
        O.lib().orc_svo_trace(O.vptr(pool), C.byref(O.gbuf_struct(g)), C.byref(p), bbmin, bbmax, C.byref(osm),
                              C.byref(O.light_struct(r["light"])), 0x3F, 0, H, O.vptr(rd), O.vptr(rs), None)
        for name, got, ref in (("diffuse", dout.cpu().numpy(), rd), ("specular", sout.cpu().numpy(), rs)):
            err = np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-12)
            print(f"mode={mode} {name}: rel L2 {err:.2e}")
            assert err <= 1e-3, (mode, name, err)


def test_empty_draw_list():
    # A mesh with no primitives: voxelization and injection leave the volumes
    # untouched, the SVO has no fragments and its pool is the 8 root children
    # (as the oracle builds it), and the context still works afterwards.
    import torch
    import vk_voxel_cone_tracing_amd as V
    from vk_voxel_cone_tracing_amd import scenes

    scene = _scene()
    pos, nrm, mat = scene.world_triangles()
    ctx = V.Context(resolution=16, levels=2, quirks=0x3F, raster_samples=1, downsample_region_size=1, extent_l0=1.0)
    ctx.update_clip_regions(np.full(3, 0.5))
    ds = V.DeviceScene(scene)
    light = scenes.light(1, scale=1.0)
    depth = O.raycast_shadow(pos, nrm, mat, light, 128, 128)
    sm = V.shadow_struct(torch.from_numpy(depth).cuda())
    ld, ls = V.light_structs(light)
    full_count = ds.mesh.primitive_count
    ds.mesh.primitive_count = 0
    ctx.clear_volume(V.VOLUME_OPACITY)
    ctx.clear_volume(V.VOLUME_RADIANCE)
    for level in range(2):
        desc = ctx.voxelization_desc(level)
        ctx.voxelize_opacity(ds, desc, level)
        ctx.inject_radiance(ds, sm, ld, ls, desc, level)
    assert not ctx.read_volume(V.VOLUME_OPACITY).any() and not ctx.read_volume(V.VOLUME_RADIANCE).any()
    nfrag = C.c_uint32(7)
    V._abi.call("vct_svo_voxelize", ctx._h, None, C.byref(ds.mesh), C.byref(sm), C.byref(ld), C.byref(ls), LEVEL,
                C.byref(nfrag))
    assert nfrag.value == 0
    V._abi.call("vct_svo_build", ctx._h, None)
    n = C.c_size_t(0)
    V._abi.call("vct_svo_read_nodes", ctx._h, None, None, 0, C.byref(n))
    nodes = np.zeros((n.value, 2), np.uint32)
    V._abi.call("vct_svo_read_nodes", ctx._h, None, nodes.ctypes.data_as(C.c_void_p), n.value, C.byref(n))
    opool = np.zeros((1000000, 2), np.uint32)
    ocount = C.c_uint32(0)
    assert O.lib().orc_svo_build(O.vptr(np.zeros((1, 2), np.uint32)), 0, LEVEL, 0x3F, O.vptr(opool), len(opool),
                                 C.byref(ocount)) == 0
    np.testing.assert_array_equal(nodes, opool[:ocount.value])
    assert len(nodes) == 8 and not nodes.any()
    # the same context with the draw list back
    ds.mesh.primitive_count = full_count
    ctx.voxelize_opacity(ds, ctx.voxelization_desc(0), 0)
    assert ctx.read_volume(V.VOLUME_OPACITY).any()


@pytest.mark.parametrize("quirks,level", [(0x3F, LEVEL), (0x00, LEVEL), (0x3F, 1)])
def test_bricks_match_dense_grid_oracle(quirks, level):
    # the 3^3 bricks of every node of the built pool against the oracle's
    # dense-grid restatement (orc_svo_bricks), bit for bit
    import vk_voxel_cone_tracing_amd as V

    print(f"[svo] bricks quirks={quirks:#x} level={level}", flush=True)
    r = _run(quirks, level=level)
    ctx = r["ctx"]
    try:
        V._abi.call("vct_svo_build_bricks", ctx._h, None)
        n = C.c_size_t(0)
        V._abi.call("vct_svo_read_bricks", ctx._h, None, None, 0, C.byref(n))
        assert n.value == len(r["nodes"])
        got = np.zeros((n.value, 27), np.uint32)
        V._abi.call("vct_svo_read_bricks", ctx._h, None, got.ctypes.data_as(C.c_void_p), n.value, C.byref(n))
    finally:
        ctx.close()
    L = O.lib()
    L.orc_svo_bricks.argtypes = [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p]
    L.orc_svo_bricks.restype = C.c_int
    ref = np.zeros_like(got)
    nodes = np.ascontiguousarray(r["nodes"])
    assert L.orc_svo_bricks(nodes.ctypes.data, len(nodes), level, ref.ctypes.data) == 0
    nonzero = int((ref != 0).any(axis=1).sum())
    print(f"{len(nodes)} nodes, {nonzero} bricks with a nonzero texel", flush=True)
    assert nonzero > 0
    np.testing.assert_array_equal(got, ref)
