"""The C ABI: libvct.so loads (without a GPU), exports every function that
include/vct.h declares, and the ctypes mirrors have the reference struct
sizes. No compute call is made here."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared_functions():
    names = set()
    for header in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(vct_[a-z0-9_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = _declared_functions()
    for required in ("vct_create", "vct_destroy", "vct_voxelize_opacity", "vct_inject_radiance", "vct_downsample",
                     "vct_wrap_border", "vct_clear_region", "vct_copy_alpha", "vct_trace"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from vk_voxel_cone_tracing_amd import _abi

    lib = C.CDLL(str(_abi.LIB_PATH))
    missing = [n for n in _declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert lib  # loaded without touching a device


def test_version_and_status_strings():
    from vk_voxel_cone_tracing_amd import _abi

    L = _abi.lib()
    assert b"gfx950" in L.vct_version()
    for code, name in _abi.STATUS_NAMES.items():
        assert L.vct_status_string(code).decode() == name


@pytest.mark.parametrize("name", sorted(__import__("vk_voxel_cone_tracing_amd._abi", fromlist=["x"]).STRUCT_SIZES))
def test_struct_sizes_match_reference_layouts(name):
    from vk_voxel_cone_tracing_amd import _abi

    cls, size = _abi.STRUCT_SIZES[name]
    assert C.sizeof(cls) == size


def test_null_arguments_are_rejected_without_a_device():
    from vk_voxel_cone_tracing_amd import _abi

    L = _abi.lib()
    assert L.vct_create(None, None) == 1
    assert L.vct_destroy(None) == 0
    assert L.vct_trace(None, None, None, None, None, None, None, None, None, None, None) == 1
    assert b"null" in L.vct_last_error()


def test_ctypes_mirror_covers_every_declared_function():
    from vk_voxel_cone_tracing_amd import _abi

    missing = [n for n in _declared_functions() if n not in _abi._PROTOS]
    assert not missing, missing


def test_ctypes_mirror_constants_match_header():
    from vk_voxel_cone_tracing_amd import _abi

    text = (ROOT / "include" / "vct.h").read_text()
    assert int(re.search(r"#define VCT_REVOX_PER_LEVEL (\d+)", text).group(1)) == _abi.REVOX_PER_LEVEL
    assert re.search(r"VCT_ADJOINT_RECOMPUTE = 1u << 0", text) and _abi.ADJOINT_RECOMPUTE == 1
    assert re.search(r"VCT_QUIRK_CLAMP_X_RANGE = 1u << 0", text) and _abi.QUIRK_CLAMP_X_RANGE == 1
