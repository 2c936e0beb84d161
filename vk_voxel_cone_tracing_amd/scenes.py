"""Seeded procedural scenes in the reference's mesh format.

The reference feeds its passes from glTF (SceneManager.cpp:124-204): separate
position / normal / texcoord streams, a u32 index buffer, per-node
(model, inverse-transpose model) matrices, GltfShadeMaterial records and one
drawIndexed per primitive with {instanceIndex, materialIndex}. These
generators produce exactly that layout from a seed (numpy PCG64, seed = the
configuration index):

* room: a unit cube of 12 inward-facing wall triangles (6 primitives), 64
  axis-aligned boxes with sides U[0.02, 0.2] (a unit box instanced with a
  scale+translate node matrix), 16 subdivision-3 icospheres (1280 triangles,
  instanced) with radius U[0.03, 0.1]; albedo U[0.2, 0.8]^3 per object;
  21 260 triangles in all;
* thin: the room plus thin triangles (edges U[0.001, 0.01]) up to a total
  triangle count (1 000 000 for the 512^3 octree workload);
* every coordinate is multiplied by `scale` (16 for the 6 x 128^3 clipmap
  workload, so level 0 spans the room);
* textured=True adds texture coordinates (walls tiled 3 x 3, box faces and
  a spherical map on the spheres) and four RGBA8 textures: a 16 x 16 and a
  5 x 3 base colour, an 8 x 8 emissive and a 32 x 32 occlusion texture whose
  red channel is below 0.1 on about a quarter of its texels. Walls, boxes and
  spheres take them in turn; the texture draws use their own generator, so
  the geometry and factors equal the untextured room's.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

MATERIAL_DTYPE = np.dtype(
    [
        ("base_color_factor", "<f4", 4),
        ("base_color_texture", "<i4"),
        ("metallic_factor", "<f4"),
        ("roughness_factor", "<f4"),
        ("metallic_roughness_texture", "<i4"),
        ("emissive_texture", "<i4"),
        ("alpha_mode", "<i4"),
        ("alpha_cutoff", "<f4"),
        ("double_sided", "<i4"),
        ("emissive_factor", "<f4", 3),
        ("normal_texture", "<i4"),
        ("normal_texture_scale", "<f4"),
        ("occlusion_texture", "<i4"),
        ("occlusion_texture_strength", "<f4"),
        ("padding", "<i4"),
    ]
)
assert MATERIAL_DTYPE.itemsize == 80


@dataclass
class Scene:
    positions: np.ndarray  # (V, 3) float32
    normals: np.ndarray  # (V, 3) float32
    texcoords: np.ndarray  # (V, 2) float32
    indices: np.ndarray  # (I,) uint32
    nodes: np.ndarray  # (N, 32) float32: model, itModel (column-major)
    materials: np.ndarray  # (M,) MATERIAL_DTYPE
    primitives: np.ndarray  # (P, 5) uint32: first_index, index_count, vertex_offset, instance, material
    surface: np.ndarray  # (M, 5) float32: perceptual roughness, metallic, specular rgb
    bb_min: np.ndarray = field(default_factory=lambda: np.zeros(3, np.float32))
    bb_max: np.ndarray = field(default_factory=lambda: np.ones(3, np.float32))
    textures: list = field(default_factory=list)  # (H, W) uint32 RGBA8 arrays, R in the low byte

    @property
    def triangle_count(self) -> int:
        return int(self.primitives[:, 1].sum() // 3)

    def triangle_uvs(self) -> np.ndarray:
        """The three vertices' texcoords of every triangle, (T, 6) float32, in
        the order of world_triangles()."""
        out = []
        for first, count, voff, inst, mat in self.primitives:
            if count == 0:
                continue
            idx = self.indices[first : first + count].astype(np.int64) + int(np.int32(voff))
            out.append(self.texcoords[idx].reshape(-1, 6))
        return np.ascontiguousarray(np.concatenate(out).astype(np.float32))

    def world_triangles(self):
        """The vertex stage (msaaVoxelizer.vert:31-37) on the host, in the
        same float32 operation order as the device kernel:
        ((m0*x + m4*y) + m8*z) + m12 and (it0*nx + it4*ny) + it8*nz."""
        pos_out, nrm_out, mat_out = [], [], []
        for first, count, voff, inst, mat in self.primitives:
            if count == 0:
                continue
            idx = self.indices[first : first + count].astype(np.int64) + int(np.int32(voff))
            p = self.positions[idx]
            n = self.normals[idx]
            m = self.nodes[inst, :16]
            it = self.nodes[inst, 16:]
            P = np.empty_like(p)
            N = np.empty_like(n)
            for r in range(3):
                P[:, r] = ((m[r] * p[:, 0] + m[4 + r] * p[:, 1]) + m[8 + r] * p[:, 2]) + m[12 + r]
                N[:, r] = (it[r] * n[:, 0] + it[4 + r] * n[:, 1]) + it[8 + r] * n[:, 2]
            pos_out.append(P.reshape(-1, 9))
            nrm_out.append(N.reshape(-1, 9))
            mat_out.append(np.full(count // 3, mat, np.uint32))
        return (
            np.ascontiguousarray(np.concatenate(pos_out).astype(np.float32)),
            np.ascontiguousarray(np.concatenate(nrm_out).astype(np.float32)),
            np.ascontiguousarray(np.concatenate(mat_out)),
        )


def _node(model: np.ndarray) -> np.ndarray:
    it = np.linalg.inv(model).T
    return np.concatenate([model.T.reshape(16), it.T.reshape(16)]).astype(np.float32)


def _material(rgb, emissive=(0.0, 0.0, 0.0)):
    m = np.zeros((), MATERIAL_DTYPE)
    m["base_color_factor"] = (rgb[0], rgb[1], rgb[2], 1.0)
    m["base_color_texture"] = -1
    m["metallic_factor"] = 0.0
    m["roughness_factor"] = 1.0
    m["metallic_roughness_texture"] = -1
    m["emissive_texture"] = -1
    m["emissive_factor"] = emissive
    m["normal_texture"] = -1
    m["normal_texture_scale"] = 1.0
    m["occlusion_texture"] = -1
    m["occlusion_texture_strength"] = 1.0
    return m


def icosphere(subdiv: int = 3):
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
         (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = list(f)
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts, np.float32)
    return V, V.copy(), np.array(faces, np.uint32).reshape(-1)


def unit_box():
    """Unit cube [-0.5, 0.5]^3, 24 vertices with flat outward normals, 12 triangles."""
    P, N, I = [], [], []
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            n = np.zeros(3)
            n[axis] = sgn
            u = np.zeros(3)
            u[(axis + 1) % 3] = 1.0
            w = np.cross(n, u)
            base = len(P)
            for cu, cw in ((-0.5, -0.5), (0.5, -0.5), (0.5, 0.5), (-0.5, 0.5)):
                P.append(n * 0.5 + u * cu + w * cw)
                N.append(n)
            I += [base, base + 1, base + 2, base, base + 2, base + 3]
    return np.array(P, np.float32), np.array(N, np.float32), np.array(I, np.uint32)


class _Builder:
    def __init__(self):
        self.P, self.N, self.I, self.nodes, self.mats, self.prims, self.surf = [], [], [], [], [], [], []
        self.UV, self.textures = [], []
        self.nv = 0
        self.ni = 0

    def mesh(self, P, N, I, UV=None):
        """Adds vertex/index data; returns (first_index, index_count, vertex_offset)."""
        first, voff = self.ni, self.nv
        self.UV.append(np.zeros((len(P), 2), np.float32) if UV is None else np.asarray(UV, np.float32))
        self.P.append(P)
        self.N.append(N)
        self.I.append(I)
        self.nv += len(P)
        self.ni += len(I)
        return first, len(I), voff

    def draw(self, mesh_ref, model, mat, surf):
        inst = len(self.nodes)
        self.nodes.append(_node(model))
        mi = len(self.mats)
        self.mats.append(mat)
        self.surf.append(surf)
        self.prims.append((mesh_ref[0], mesh_ref[1], mesh_ref[2], inst, mi))

    def build(self, scale) -> Scene:
        P = np.concatenate(self.P).astype(np.float32)
        lo, hi = np.full(3, np.inf), np.full(3, -np.inf)
        for (first, count, voff, inst, mat) in self.prims:
            idx = np.concatenate(self.I)[first : first + count].astype(np.int64) + voff
            m = self.nodes[inst][:16].reshape(4, 4).T.astype(np.float64)
            w = (m[:3, :3] @ P[idx].T.astype(np.float64)).T + m[:3, 3]
            lo = np.minimum(lo, w.min(0))
            hi = np.maximum(hi, w.max(0))
        return Scene(
            positions=P,
            normals=np.concatenate(self.N).astype(np.float32),
            texcoords=np.concatenate(self.UV).astype(np.float32),
            indices=np.concatenate(self.I).astype(np.uint32),
            nodes=np.stack(self.nodes).astype(np.float32),
            materials=np.array(self.mats, MATERIAL_DTYPE),
            primitives=np.array(self.prims, np.int64).astype(np.uint32),
            surface=np.array(self.surf, np.float32),
            bb_min=lo.astype(np.float32),
            bb_max=hi.astype(np.float32),
            textures=list(self.textures),
        )


def _surface(rng):
    # perceptual roughness, metallic, specular rgb (G-buffer material inputs)
    return (rng.uniform(0.1, 0.8), rng.uniform(0.1, 0.9), *rng.uniform(0.2, 0.8, 3))


def _texture(rng, h, w, lo=0, hi=256):
    t = rng.integers(lo, hi, (h, w, 4)).astype(np.uint32)
    return np.ascontiguousarray(t[..., 0] | (t[..., 1] << 8) | (t[..., 2] << 16) | (t[..., 3] << 24))


def _room_textures(seed: int):
    """Base colour 16 x 16 and 5 x 3 (opaque), emissive 8 x 8, occlusion
    32 x 32 (red below 0.1 on about a quarter of the texels)."""
    rng = np.random.Generator(np.random.PCG64(5000 + seed))
    base0 = _texture(rng, 16, 16) | np.uint32(0xFF000000)
    base1 = _texture(rng, 3, 5) | np.uint32(0xFF000000)
    emis = _texture(rng, 8, 8)
    red = np.where(rng.uniform(size=(32, 32)) < 0.25, rng.integers(0, 20, (32, 32)), rng.integers(40, 256, (32, 32)))
    occ = (_texture(rng, 32, 32) & np.uint32(0xFFFFFF00)) | red.astype(np.uint32)
    return [base0, base1, emis, occ]


def _with_textures(m, base=-1, emissive=-1, occlusion=-1):
    m = m.copy()
    m["base_color_texture"] = base
    m["emissive_texture"] = emissive
    m["occlusion_texture"] = occlusion
    return m


def room(seed: int, scale: float = 1.0, n_boxes: int = 64, n_spheres: int = 16, thin_total: int = 0,
         emissive_boxes: int = 0, textured: bool = False) -> Scene:
    rng = np.random.Generator(np.random.PCG64(seed))
    S = np.diag([scale, scale, scale, 1.0])
    b = _Builder()
    if textured:
        b.textures = _room_textures(seed)
    # texture indices by object number: (base colour, emissive, occlusion)
    tex_of = lambda k, emissive: (-1 if not textured else (
        (-1, 2 if emissive else -1, 3) if k % 4 == 3 else (k % 2, 2 if emissive else -1, 3 if k % 3 == 0 else -1)))
    # 12 inward-facing wall triangles, one primitive per wall
    for axis in range(3):
        for side in (0.0, 1.0):
            n = np.zeros(3)
            n[axis] = 1.0 if side == 0.0 else -1.0  # inward
            u = np.zeros(3)
            u[(axis + 1) % 3] = 1.0
            w = np.cross(n, u)
            c = np.full(3, 0.5)
            c[axis] = side
            quad = [c + u * cu + w * cw for cu, cw in ((-0.5, -0.5), (0.5, -0.5), (0.5, 0.5), (-0.5, 0.5))]
            P = np.array(quad, np.float32)
            N = np.tile(n.astype(np.float32), (4, 1))
            UV = np.array([(0.0, 0.0), (3.0, 0.0), (3.0, 3.0), (0.0, 3.0)], np.float32)  # tiled 3 x 3 (REPEAT)
            ref = b.mesh(P, N, np.array([0, 1, 2, 0, 2, 3], np.uint32), UV)
            mat = _material(rng.uniform(0.2, 0.8, 3))
            if textured:
                mat = _with_textures(mat, *tex_of(2 * axis + int(side), False))
            b.draw(ref, S, mat, _surface(rng))
    bp, bn, bi = unit_box()
    box_uv = np.tile(np.array([(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)], np.float32), (6, 1))
    box_ref = b.mesh(bp, bn, bi, box_uv)
    for k in range(n_boxes):
        size = rng.uniform(0.02, 0.2, 3)
        center = rng.uniform(size / 2, 1.0 - size / 2)
        model = S @ np.array([[size[0], 0, 0, center[0]], [0, size[1], 0, center[1]],
                              [0, 0, size[2], center[2]], [0, 0, 0, 1.0]])
        em = tuple(rng.uniform(0.5, 1.0, 3)) if k < emissive_boxes else (0.0, 0.0, 0.0)
        mat = _material(rng.uniform(0.2, 0.8, 3), em)
        if textured:
            mat = _with_textures(mat, *tex_of(k, k < emissive_boxes))
        b.draw(box_ref, model, mat, _surface(rng))
    sp, sn, si = icosphere(3)
    sphere_uv = np.stack([np.arctan2(sp[:, 2], sp[:, 0]) / (2 * np.pi) + 0.5,
                          np.arccos(np.clip(sp[:, 1], -1.0, 1.0)) / np.pi], axis=1).astype(np.float32)
    sphere_ref = b.mesh(sp, sn, si, sphere_uv)
    for _ in range(n_spheres):
        r = rng.uniform(0.03, 0.1)
        center = rng.uniform(r, 1.0 - r, 3)
        model = S @ np.array([[r, 0, 0, center[0]], [0, r, 0, center[1]], [0, 0, r, center[2]], [0, 0, 0, 1.0]])
        mat = _material(rng.uniform(0.2, 0.8, 3))
        if textured:
            mat = _with_textures(mat, *tex_of(len(b.prims), False))
        b.draw(sphere_ref, model, mat, _surface(rng))
    room_tris = 12 + n_boxes * 12 + n_spheres * 1280
    n_thin = max(0, thin_total - room_tris)
    if n_thin:
        p0 = rng.uniform(0.0, 1.0, (n_thin, 3))
        d1 = rng.normal(size=(n_thin, 3))
        d2 = rng.normal(size=(n_thin, 3))
        d1 /= np.linalg.norm(d1, axis=1, keepdims=True)
        d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
        e1 = rng.uniform(0.001, 0.01, (n_thin, 1))
        e2 = rng.uniform(0.001, 0.01, (n_thin, 1))
        tri = np.stack([p0, p0 + d1 * e1, p0 + d2 * e2], axis=1)  # (n, 3, 3)
        nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
        nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
        P = tri.reshape(-1, 3).astype(np.float32)
        N = np.repeat(nrm, 3, axis=0).astype(np.float32)
        groups = 8
        per = -(-n_thin // groups)
        for g in range(groups):
            lo, hi = g * per, min(n_thin, (g + 1) * per)
            if lo >= hi:
                break
            ref = b.mesh(P[3 * lo : 3 * hi], N[3 * lo : 3 * hi], np.arange(3 * (hi - lo), dtype=np.uint32))
            b.draw(ref, S, _material(rng.uniform(0.2, 0.8, 3)), _surface(rng))
    return b.build(scale)


def camera(seed: int, width: int, height: int, scale: float = 1.0, fovy_deg: float = 60.0):
    """Seeded camera inside the room: (view_proj, view_proj_inv, eye), glm LH_ZO
    perspective with the reference's near/far (Camera.cpp:112-113)."""
    from . import glm

    rng = np.random.Generator(np.random.PCG64(1000 + seed))
    eye = rng.uniform(0.3, 0.7, 3) * scale
    yaw = rng.uniform(0, 2 * np.pi)
    pitch = rng.uniform(-0.3, 0.3)
    d = np.array([np.cos(pitch) * np.cos(yaw), np.sin(pitch), np.cos(pitch) * np.sin(yaw)])
    view = glm.look_at_lh(eye, eye + d, np.array([0.0, 1.0, 0.0]))
    proj = glm.perspective_lh_zo(np.radians(fovy_deg), width / height, 0.01, 5000.0)
    vp = proj @ view
    vp_inv = np.linalg.inv(view) @ np.linalg.inv(proj)
    return vp, vp_inv, eye


def light(seed: int, scale: float = 1.0):
    """Seeded directional light looking down into the room, with the
    reference's orthographic shadow projection (DirectionalLight.cpp:21-47):
    lookAt(origin, origin + dir, +y) and ortho(+-half, near 0.1, far)."""
    from . import glm

    rng = np.random.Generator(np.random.PCG64(2000 + seed))
    d = np.array([rng.uniform(-0.4, 0.4), -1.0, rng.uniform(-0.4, 0.4)])
    d /= np.linalg.norm(d)
    center = np.full(3, 0.5 * scale)
    origin = center - d * 1.5 * scale
    half = 0.75 * scale
    view = glm.look_at_lh(origin, origin + d, np.array([0.0, 1.0, 0.0]))
    proj = glm.ortho_lh_zo(-half, half, -half, half, 0.1, 3.0 * scale)
    return dict(direction=d.astype(np.float32), color=np.ones(3, np.float32), intensity=np.float32(1.0),
                view=view, proj=proj, z_near=np.float32(0.1), z_far=np.float32(3.0 * scale))


def batch(seed: int = 3, count: int = 64, tri_min: int = 5_000, tri_max: int = 500_000):
    """Scenes of uneven size for the batched workload: triangle counts
    log-uniform in [tri_min, tri_max] and dense grid resolutions log-uniform
    over 64^3..256^3 (the exponent drawn from U[6, 8] and rounded, since the
    mip pyramid needs a power of two). Each scene is the room with as many of
    its 16 spheres as fit the budget, topped up with thin triangles.
    Returns a list of (Scene, resolution)."""
    out = []
    room_fixed = 12 + 64 * 12
    for i, (tris, res) in enumerate(batch_sizes(seed, count, tri_min, tri_max)):
        spheres = int(min(16, max(0, (tris - room_fixed) // 1280)))
        scene = room(10_000 + 100 * seed + i, scale=1.0, n_boxes=64, n_spheres=spheres, thin_total=tris)
        out.append((scene, res))
    return out


def batch_sizes(seed: int = 3, count: int = 64, tri_min: int = 5_000, tri_max: int = 500_000):
    """The (triangle count, grid resolution) pairs of batch() without building
    the geometry."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        tris = int(round(np.exp(rng.uniform(np.log(tri_min), np.log(tri_max)))))
        res = 1 << int(round(rng.uniform(6.0, 8.0)))
        out.append((tris, res))
    return out
