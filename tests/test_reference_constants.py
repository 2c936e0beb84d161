"""The reference's own constants (tests/golden/reference_constants.json, the
numbers its shaders and passes hold) against what the oracle and the host
layer use: cone tables and apertures, PCF lattice, trace defaults, clip-region
snapping, engine configuration."""
import ctypes as C
import json
from pathlib import Path

import numpy as np

import oracle_api as O

FX = json.loads((Path(__file__).parent / "golden" / "reference_constants.json").read_text())


def _oracle_dirs(n):
    L = O.lib()
    fn = L.orc_dense_dir16 if n == 16 else L.orc_dense_dir32
    fn.restype = C.POINTER(C.c_double)
    fn.argtypes = [C.c_int]
    return np.array([[fn(i)[k] for k in range(3)] for i in range(n)])


def test_cone_tables_match_reference():
    for n, key in ((16, "cone_directions_16"), (32, "cone_directions_32")):
        ref = np.array(FX[key]["dirs"])
        np.testing.assert_array_equal(_oracle_dirs(n), ref)
        # the tables are unit directions to the precision they are printed with
        assert np.abs(np.linalg.norm(ref, axis=1) - 1.0).max() < 2e-5


def test_trace_defaults_match_reference():
    from vk_voxel_cone_tracing_amd import passes

    ref = FX["cone_tracing_defaults"]
    p = passes.VoxelConeTracingPass.__new__(passes.VoxelConeTracingPass)
    passes.VoxelConeTracingPass.__init__(p, None, None, None, None, None, None, None, None)
    for k in ("trace_start_offset", "indirect_diffuse_intensity", "ambient_occlusion_factor", "min_trace_step_factor",
              "indirect_specular_intensity", "occlusion_decay"):
        assert getattr(p, k) == ref[k], k
    assert int(p.enable_32_cones) == ref["enable_32_cones"]


def test_engine_config_defaults_match_reference():
    import inspect

    import vk_voxel_cone_tracing_amd as V

    ref = FX["engine_config"]
    sig = inspect.signature(V.Context.__init__)
    assert sig.parameters["resolution"].default == ref["voxel_resolution"]
    assert sig.parameters["levels"].default == ref["clip_region_count"]


def test_pcf_lattice_is_the_reference_one():
    # orc_visibility samples the 4x4 lattice of offsets {-1.5..1.5} texels:
    # on a depth map that is a linear ramp in x, the raw-depth average equals
    # the ramp at the centre (the lattice is symmetric about it)
    offs = FX["pcf_offsets"]["texel_offsets"]
    assert offs == [-1.5, -0.5, 0.5, 1.5] and abs(sum(offs)) == 0.0
