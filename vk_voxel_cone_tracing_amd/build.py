"""Builds libvct.so (the gfx950 HIP library) in-tree with hipcc, and the
libtorch binding _vct_torch*.so of the dense adjoint on top of it.

Usage: python -m vk_voxel_cone_tracing_amd.build  (or build_all() from Python).
Every .hip file under csrc/ is compiled to an object in parallel and linked
into vk_voxel_cone_tracing_amd/libvct.so; csrc_torch/vct_torch.cpp is
compiled against PyTorch's headers and linked to libvct.so (rpath $ORIGIN).
The build always recompiles.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libvct.so"

ARCH = os.environ.get("VCT_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 ops (v_pk_*_f32) are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-result",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: Path) -> Path:
    obj = OBJ / (src.stem + ".o")
    cmd = [_hipcc(), *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")
    return obj


def build_all(verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    workers = min(len(sources), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(sources)} sources", file=sys.stderr)
    return LIB


TORCH_SRC = PKG / "csrc_torch" / "vct_torch.cpp"


def torch_ext_path() -> Path:
    import sysconfig

    return PKG / ("_vct_torch" + sysconfig.get_config_var("EXT_SUFFIX"))


def build_torch_ext(verbose: bool = True) -> Path:
    """The autograd binding: host C++ only (no device code), compiled with
    hipcc so the HIP runtime headers that ATen's HIP context uses resolve."""
    import sysconfig

    import torch
    import torch.utils.cpp_extension as ce

    out = torch_ext_path()
    tmp = out.with_name(out.name + ".tmp")
    incs = [f"-I{p}" for p in ce.include_paths(device_type="cuda")] + [f"-I{sysconfig.get_paths()['include']}"]
    libdirs = ce.library_paths(device_type="cuda")
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cmd = [
        _hipcc(), "-O2", "-std=c++17", "-fPIC", "-shared", "-w",
        "-DTORCH_EXTENSION_NAME=_vct_torch", "-DTORCH_API_INCLUDE_EXTENSION_H",
        f"-D_GLIBCXX_USE_CXX11_ABI={abi}", "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1",
        *incs, str(TORCH_SRC), "-o", str(tmp),
        *[f"-L{d}" for d in libdirs], "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_hip",
        "-ltorch_python", f"-L{PKG}", "-lvct", "-Wl,-rpath,$ORIGIN",
        *[f"-Wl,-rpath,{d}" for d in libdirs],
    ]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"torch binding build failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}", file=sys.stderr)
    return out


if __name__ == "__main__":
    build_all()
    build_torch_ext()
