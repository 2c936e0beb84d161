"""ctypes mirror of include/vct.h (the C ABI of libvct.so).

The library is loaded from the package directory (built in-tree by
vk_voxel_cone_tracing_amd.build). Loading fails loudly if it is missing: there
is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libvct.so"

VCT_OK = 0
STATUS_NAMES = {
    0: "VCT_OK",
    1: "VCT_ERR_INVALID_ARGUMENT",
    2: "VCT_ERR_OUT_OF_MEMORY",
    3: "VCT_ERR_HIP",
    4: "VCT_ERR_CAPACITY",
    5: "VCT_ERR_UNSUPPORTED",
    6: "VCT_ERR_NOT_READY",
}
for _code, _name in STATUS_NAMES.items():
    globals()[_name] = _code

VOLUME_OPACITY = 0
VOLUME_RADIANCE = 1
FORMAT_RGBA8 = 0
FORMAT_RGBA16F = 1

QUIRK_CLAMP_X_RANGE = 1 << 0
QUIRK_BORDER_SKIP_FAR = 1 << 1
QUIRK_SHADOW_RAW_DEPTH = 1 << 2
QUIRK_SVO_HALF_COORDS = 1 << 3
QUIRK_SVO_MIP_LEVEL_TEST = 1 << 4
QUIRK_SVO_UNTEXTURED_BLACK = 1 << 5
QUIRK_ALL = 0x3F
ADJOINT_RECOMPUTE = 1
REVOX_PER_LEVEL = 4  # VCT_REVOX_PER_LEVEL: regions per level written by vct_move_clip_regions
BATCH_PHASES = 12  # VCT_BATCH_PHASES: need[] entries per scene written by vct_batch_plan

MODE_DIFFUSE = 0
MODE_SPECULAR = 1
MODE_NORMAL = 2
MODE_MIN_LEVEL = 3
MODE_DIRECT = 4
MODE_INDIRECT_DIFFUSE = 5
MODE_INDIRECT_SPECULAR = 6
MODE_AMBIENT_OCCLUSION = 7
MODE_COMBINED_GI = 8

f3 = C.c_float * 3
i3 = C.c_int32 * 3
u3 = C.c_uint32 * 3
m4 = C.c_float * 16


class Config(C.Structure):
    _fields_ = [
        ("voxel_resolution", C.c_uint32),
        ("clip_levels", C.c_uint32),
        ("radiance_format", C.c_uint32),
        ("quirks", C.c_uint32),
        ("raster_samples", C.c_uint32),
        ("downsample_region_size", C.c_uint32),
        ("extent_l0", C.c_float),
        ("reserved", C.c_uint32),
    ]


class ClipmapRegion(C.Structure):
    _fields_ = [("min_corner", i3), ("extent", u3), ("voxel_size", C.c_float)]


class VoxelizationDesc(C.Structure):
    _fields_ = [
        ("region_min_corner", f3),
        ("clip_level", C.c_uint32),
        ("region_max_corner", f3),
        ("clip_max_extent", C.c_float),
        ("prev_region_min_corner", f3),
        ("voxel_size", C.c_float),
        ("prev_region_max_corner", f3),
        ("clipmap_resolution", C.c_int32),
    ]


class DownSampleDesc(C.Structure):
    _fields_ = [
        ("prev_region_min_corner", i3),
        ("clip_level", C.c_int32),
        ("clipmap_resolution", C.c_int32),
        ("downsample_region_size", C.c_int32),
    ]


class ImageCleaningDesc(C.Structure):
    _fields_ = [
        ("region_min_corner", i3),
        ("clip_level", C.c_int32),
        ("clip_max_extent", u3),
        ("clipmap_resolution", C.c_int32),
        ("face_count", C.c_uint32),
    ]


class BorderWrappingDesc(C.Structure):
    _fields_ = [
        ("clipmap_resolution", C.c_uint32),
        ("clip_border_width", C.c_uint32),
        ("face_count", C.c_uint32),
        ("clip_region_count", C.c_uint32),
    ]


class CopyAlphaDesc(C.Structure):
    _fields_ = [("clip_level", C.c_int32), ("clipmap_resolution", C.c_int32), ("face_count", C.c_int32)]


class ConeTracingDesc(C.Structure):
    _fields_ = [
        ("volume_center", f3),
        ("rendering_mode", C.c_uint32),
        ("voxel_size", C.c_float),
        ("volume_dimension", C.c_float),
        ("trace_start_offset", C.c_float),
        ("indirect_diffuse_intensity", C.c_float),
        ("ambient_occlusion_factor", C.c_float),
        ("min_trace_step_factor", C.c_float),
        ("indirect_specular_intensity", C.c_float),
        ("occlusion_decay", C.c_float),
        ("enable_32_cones", C.c_int32),
    ]


class DirectionalLightDesc(C.Structure):
    _fields_ = [("direction", f3), ("intensity", C.c_float), ("color", f3), ("padding", C.c_int32)]


class DirectionalLightShadowDesc(C.Structure):
    _fields_ = [("view", m4), ("proj", m4), ("z_near", C.c_float), ("z_far", C.c_float), ("padding", C.c_float * 2)]


class CameraUBO(C.Structure):
    _fields_ = [("view_proj", m4), ("view_proj_inv", m4), ("eye_pos", f3), ("padding", C.c_int32)]


class GltfMaterial(C.Structure):
    _fields_ = [
        ("base_color_factor", C.c_float * 4),
        ("base_color_texture", C.c_int32),
        ("metallic_factor", C.c_float),
        ("roughness_factor", C.c_float),
        ("metallic_roughness_texture", C.c_int32),
        ("emissive_texture", C.c_int32),
        ("alpha_mode", C.c_int32),
        ("alpha_cutoff", C.c_float),
        ("double_sided", C.c_int32),
        ("emissive_factor", f3),
        ("normal_texture", C.c_int32),
        ("normal_texture_scale", C.c_float),
        ("occlusion_texture", C.c_int32),
        ("occlusion_texture_strength", C.c_float),
        ("padding", C.c_int32),
    ]


class Primitive(C.Structure):
    _fields_ = [
        ("first_index", C.c_uint32),
        ("index_count", C.c_uint32),
        ("vertex_offset", C.c_int32),
        ("instance_index", C.c_uint32),
        ("material_index", C.c_uint32),
    ]


class Texture(C.Structure):
    _fields_ = [("texels", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32)]


class Mesh(C.Structure):
    _fields_ = [
        ("positions", C.c_void_p),
        ("normals", C.c_void_p),
        ("texcoords", C.c_void_p),
        ("indices", C.c_void_p),
        ("node_matrices", C.c_void_p),
        ("materials", C.c_void_p),
        ("vertex_count", C.c_uint32),
        ("index_count", C.c_uint32),
        ("instance_count", C.c_uint32),
        ("material_count", C.c_uint32),
        ("primitives", C.POINTER(Primitive)),
        ("primitive_count", C.c_uint32),
        ("scene_bb_min", f3),
        ("scene_bb_max", f3),
        ("textures", C.POINTER(Texture)),
        ("texture_count", C.c_uint32),
        ("material_version", C.c_uint32),
    ]


class ShadowMap(C.Structure):
    _fields_ = [("depth", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32)]


class GBuffer(C.Structure):
    _fields_ = [
        ("diffuse", C.c_void_p),
        ("normal", C.c_void_p),
        ("specular", C.c_void_p),
        ("emission", C.c_void_p),
        ("depth", C.c_void_p),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
    ]


class DenseDesc(C.Structure):
    _fields_ = [
        ("grid_min", f3),
        ("extent", C.c_float),
        ("resolution", C.c_uint32),
        ("format", C.c_uint32),
        ("raster_samples", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class DenseTraceDesc(C.Structure):
    _fields_ = [
        ("aperture", C.c_float),
        ("max_distance", C.c_float),
        ("min_trace_step_factor", C.c_float),
        ("trace_start_offset", C.c_float),
        ("occlusion_decay", C.c_float),
        ("ambient_occlusion_factor", C.c_float),
        ("indirect_diffuse_intensity", C.c_float),
        ("enable_32_cones", C.c_int32),
    ]


class SpecularFilterDesc(C.Structure):
    _fields_ = [
        ("tonemap_gamma", C.c_float),
        ("tonemap_exposure", C.c_float),
        ("tonemap_enable", C.c_int32),
        ("filter_method", C.c_int32),
    ]


class BatchScene(C.Structure):
    _fields_ = [
        ("mesh", Mesh),
        ("grid", DenseDesc),
        ("gbuffer", GBuffer),
        ("camera", CameraUBO),
        ("out", C.c_void_p),
    ]


class BatchTuning(C.Structure):
    _fields_ = [
        ("raster_chunk", C.c_uint32),
        ("mip_brick", C.c_uint32),
        ("trace_schedule", C.c_uint32),
        ("workgroups_per_cu", C.c_uint32),
    ]


class BatchSceneShape(C.Structure):
    _fields_ = [
        ("triangles", C.c_uint32),
        ("resolution", C.c_uint32),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
    ]


STRUCT_SIZES = {
    "vct_config": (Config, 32),
    "vct_clipmap_region": (ClipmapRegion, 28),
    "vct_voxelization_desc": (VoxelizationDesc, 64),
    "vct_downsample_desc": (DownSampleDesc, 24),
    "vct_image_cleaning_desc": (ImageCleaningDesc, 36),
    "vct_border_wrapping_desc": (BorderWrappingDesc, 16),
    "vct_copy_alpha_desc": (CopyAlphaDesc, 12),
    "vct_cone_tracing_desc": (ConeTracingDesc, 52),
    "vct_directional_light_desc": (DirectionalLightDesc, 32),
    "vct_directional_light_shadow_desc": (DirectionalLightShadowDesc, 144),
    "vct_camera_ubo": (CameraUBO, 144),
    "vct_gltf_material": (GltfMaterial, 80),
    "vct_dense_desc": (DenseDesc, 32),
    "vct_dense_trace_desc": (DenseTraceDesc, 32),
    "vct_mesh": (Mesh, 120),
    "vct_texture": (Texture, 16),
    "vct_gbuffer": (GBuffer, 48),
    "vct_batch_scene": (BatchScene, 352),
    "vct_batch_tuning": (BatchTuning, 16),
    "vct_batch_scene_shape": (BatchSceneShape, 16),
    "vct_specular_filter_desc": (SpecularFilterDesc, 16),
}

P = C.c_void_p
_PROTOS = {
    "vct_version": (C.c_char_p, []),
    "vct_status_string": (C.c_char_p, [C.c_int]),
    "vct_last_error": (C.c_char_p, []),
    "vct_create": (C.c_int, [C.POINTER(P), C.POINTER(Config)]),
    "vct_destroy": (C.c_int, [P]),
    "vct_set_clip_regions": (C.c_int, [P, C.POINTER(ClipmapRegion), C.c_uint32]),
    "vct_get_clip_regions": (C.c_int, [P, C.POINTER(ClipmapRegion), C.c_uint32]),
    "vct_update_clip_regions": (C.c_int, [P, C.POINTER(C.c_float)]),
    "vct_compute_clip_regions": (C.c_int, [C.POINTER(Config), C.POINTER(C.c_float), C.POINTER(ClipmapRegion),
                                           C.c_uint32]),
    "vct_move_clip_regions": (C.c_int, [P, C.POINTER(C.c_float), C.POINTER(ClipmapRegion), C.POINTER(C.c_uint32)]),
    "vct_make_voxelization_desc": (C.c_int, [P, C.c_uint32, C.POINTER(ClipmapRegion), C.POINTER(VoxelizationDesc)]),
    "vct_clear_volume": (C.c_int, [P, P, C.c_int]),
    "vct_voxelize_opacity": (C.c_int, [P, P, C.POINTER(Mesh), C.POINTER(VoxelizationDesc), C.c_uint32]),
    "vct_inject_radiance": (
        C.c_int,
        [P, P, C.POINTER(Mesh), C.POINTER(ShadowMap), C.POINTER(DirectionalLightDesc),
         C.POINTER(DirectionalLightShadowDesc), C.POINTER(VoxelizationDesc), C.c_uint32],
    ),
    "vct_voxelize_opacity_levels": (C.c_int, [P, P, C.POINTER(Mesh), C.POINTER(VoxelizationDesc), C.c_uint32]),
    "vct_inject_radiance_levels": (
        C.c_int,
        [P, P, C.POINTER(Mesh), C.POINTER(ShadowMap), C.POINTER(DirectionalLightDesc),
         C.POINTER(DirectionalLightShadowDesc), C.POINTER(VoxelizationDesc), C.c_uint32],
    ),
    "vct_clear_region": (C.c_int, [P, P, C.c_int, C.POINTER(ImageCleaningDesc)]),
    "vct_downsample": (C.c_int, [P, P, C.c_int, C.POINTER(DownSampleDesc)]),
    "vct_copy_alpha": (C.c_int, [P, P, C.POINTER(CopyAlphaDesc)]),
    "vct_begin_radiance_level": (C.c_int, [P, P, C.c_uint32]),
    "vct_wrap_border": (C.c_int, [P, P, C.c_int, C.POINTER(BorderWrappingDesc)]),
    "vct_trace": (
        C.c_int,
        [P, P, C.POINTER(GBuffer), C.POINTER(CameraUBO), C.POINTER(ConeTracingDesc), C.POINTER(ShadowMap),
         C.POINTER(DirectionalLightDesc), C.POINTER(DirectionalLightShadowDesc), P, P, P],
    ),
    "vct_trace_noalloc": (
        C.c_int,
        [P, P, C.POINTER(GBuffer), C.POINTER(CameraUBO), C.POINTER(ConeTracingDesc), C.POINTER(ShadowMap),
         C.POINTER(DirectionalLightDesc), C.POINTER(DirectionalLightShadowDesc), P, P, P],
    ),
    "vct_raycast_gbuffer": (
        C.c_int, [P, P, C.POINTER(Mesh), P, C.POINTER(CameraUBO), P, P, P, P, P, C.c_uint32, C.c_uint32]
    ),
    "vct_raycast_shadow": (C.c_int, [P, P, C.POINTER(Mesh), C.POINTER(DirectionalLightShadowDesc), P, C.c_uint32, C.c_uint32]),
    "vct_svo_voxelize": (
        C.c_int,
        [P, P, C.POINTER(Mesh), C.POINTER(ShadowMap), C.POINTER(DirectionalLightDesc),
         C.POINTER(DirectionalLightShadowDesc), C.c_uint32, C.POINTER(C.c_uint32)],
    ),
    "vct_svo_build": (C.c_int, [P, P]),
    "vct_svo_read_nodes": (C.c_int, [P, P, P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "vct_svo_read_fragments": (C.c_int, [P, P, P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "vct_svo_pool": (C.c_int, [P, C.POINTER(C.c_void_p), C.POINTER(C.c_uint32)]),
    "vct_svo_build_bricks": (C.c_int, [P, P]),
    "vct_svo_read_bricks": (C.c_int, [P, P, P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "vct_svo_trace": (
        C.c_int,
        [P, P, C.POINTER(GBuffer), C.POINTER(CameraUBO), C.POINTER(ConeTracingDesc), C.POINTER(ShadowMap),
         C.POINTER(DirectionalLightDesc), C.POINTER(DirectionalLightShadowDesc), C.POINTER(C.c_float),
         C.POINTER(C.c_float), P, P, P],
    ),
    "vct_dense_create": (C.c_int, [C.POINTER(P), C.POINTER(DenseDesc)]),
    "vct_dense_destroy": (C.c_int, [P]),
    "vct_dense_clear": (C.c_int, [P, P]),
    "vct_dense_level": (C.c_int, [P, C.c_uint32, C.POINTER(C.c_void_p), C.POINTER(C.c_size_t), C.POINTER(C.c_uint32)]),
    "vct_dense_read_level": (C.c_int, [P, P, C.c_uint32, P, C.c_size_t]),
    "vct_dense_write_level": (C.c_int, [P, P, C.c_uint32, P, C.c_size_t]),
    "vct_dense_voxelize": (C.c_int, [P, P, C.POINTER(Mesh)]),
    "vct_dense_mipmap": (C.c_int, [P, P]),
    "vct_dense_trace": (
        C.c_int, [P, P, C.POINTER(GBuffer), C.POINTER(CameraUBO), C.POINTER(DenseTraceDesc), P, P]
    ),
    "vct_dense_texels": (C.c_int, [P, C.POINTER(C.c_size_t)]),
    "vct_dense_set_level0": (C.c_int, [P, P, P, P]),
    "vct_dense_trace_backward": (
        C.c_int, [P, P, C.POINTER(GBuffer), C.POINTER(CameraUBO), C.POINTER(DenseTraceDesc), P, P, C.c_uint32, P]
    ),
    "vct_dense_tape_bytes": (
        C.c_int, [P, C.POINTER(GBuffer), C.POINTER(CameraUBO), C.POINTER(DenseTraceDesc), C.POINTER(C.c_size_t)]
    ),
    "vct_dense_trace_tape": (
        C.c_int, [P, P, C.POINTER(GBuffer), C.POINTER(CameraUBO), C.POINTER(DenseTraceDesc), P, P, P, C.c_size_t]
    ),
    "vct_dense_trace_backward_tape": (
        C.c_int, [P, P, C.POINTER(GBuffer), C.POINTER(CameraUBO), C.POINTER(DenseTraceDesc), P, P, C.c_size_t, P, P]
    ),
    "vct_dense_trace_backward_tapes": (
        C.c_int, [P, P, C.c_uint32, C.POINTER(GBuffer), C.POINTER(CameraUBO), C.POINTER(DenseTraceDesc),
                  C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_size_t, P, P]
    ),
    "vct_dense_mipmap_backward": (C.c_int, [P, P, P]),
    "vct_dense_level0_grad": (C.c_int, [P, P, P, P, P]),
    "vct_batch_create": (C.c_int, [C.POINTER(P), C.POINTER(BatchScene), C.c_uint32, C.POINTER(DenseTraceDesc)]),
    "vct_batch_destroy": (C.c_int, [P]),
    "vct_batch_set_tuning": (C.c_int, [P, C.POINTER(BatchTuning)]),
    "vct_batch_get_tuning": (C.c_int, [P, C.POINTER(BatchTuning)]),
    "vct_batch_run": (C.c_int, [P, P, P]),
    "vct_batch_status": (C.c_int, [P]),
    "vct_batch_autotune": (C.c_int, [P, P, C.c_uint32, C.POINTER(BatchTuning), C.POINTER(C.c_float)]),
    "vct_batch_level": (C.c_int, [P, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p), C.POINTER(C.c_size_t),
                                  C.POINTER(C.c_uint32)]),
    "vct_batch_read_level": (C.c_int, [P, P, C.c_uint32, C.c_uint32, P, C.c_size_t]),
    "vct_batch_plan": (C.c_int, [C.POINTER(BatchSceneShape), C.c_uint32, C.POINTER(BatchTuning), P, C.c_size_t,
                                 C.POINTER(C.c_size_t), P]),
    "vct_specular_filter": (C.c_int, [P, P, P, C.c_uint32, C.c_uint32, C.POINTER(SpecularFilterDesc), P]),
    "vct_volume_bytes": (C.c_int, [P, C.c_int, C.POINTER(C.c_size_t)]),
    "vct_read_volume": (C.c_int, [P, P, C.c_int, P, C.c_size_t]),
    "vct_write_volume": (C.c_int, [P, P, C.c_int, P, C.c_size_t]),
}


class VctError(RuntimeError):
    def __init__(self, fn: str, status: int, detail: str):
        super().__init__(f"{fn} -> {STATUS_NAMES.get(status, status)}: {detail}")
        self.status = status


_lib = None


def lib() -> C.CDLL:
    """Loads libvct.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m vk_voxel_cone_tracing_amd.build` "
                "(there is no CPU fallback)"
            )
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _PROTOS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(fn: str, status: int) -> None:
    if status != VCT_OK:
        detail = lib().vct_last_error().decode(errors="replace")
        raise VctError(fn, status, detail)


def call(fn: str, *args) -> None:
    check(fn, getattr(lib(), fn)(*args))
