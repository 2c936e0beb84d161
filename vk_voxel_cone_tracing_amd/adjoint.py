"""Differentiable dense render: the forward VXAO / indirect-diffuse trace of an
RGBA32F dense grid and its hand-written adjoint, as a torch autograd
function implemented in C++ (csrc_torch/vct_torch.cpp, a
torch::autograd::Function over libvct's vct_dense_* entry points).

The forward model is the dense method (vct_dense_mipmap then
vct_dense_trace) applied to a level-0 grid of albedo (rgb) and opacity (a);
the backward returns gradients with respect to both. There is no CPU path:
importing this module fails if the binding has not been built.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import build as _build

if not _build.torch_ext_path().exists():
    raise ImportError(f"{_build.torch_ext_path()} is missing: build it with `python -m vk_voxel_cone_tracing_amd.build`")

import torch  # noqa: E402

from . import _vct_torch  # noqa: E402


@dataclass
class TraceParams:
    """vct_dense_trace_desc: the reference's VoxelConeTracingDesc defaults
    (VoxelConeTracingPass.h:75-82) with the 16-cone aperture."""

    aperture: float = 0.872665
    max_distance: float = 1.7320508
    min_trace_step_factor: float = 1.0
    trace_start_offset: float = 1.0
    occlusion_decay: float = 2.0
    ambient_occlusion_factor: float = 0.5
    indirect_diffuse_intensity: float = 8.0
    enable_32_cones: int = 0

    def as_list(self):
        return [float(self.aperture), float(self.max_distance), float(self.min_trace_step_factor),
                float(self.trace_start_offset), float(self.occlusion_decay), float(self.ambient_occlusion_factor),
                float(self.indirect_diffuse_intensity), float(self.enable_32_cones)]


def cameras_tensor(cams) -> torch.Tensor:
    """[V, 36] float32 host tensor from (view_proj, view_proj_inv, eye) triples
    (glm column-major)."""
    from . import glm

    rows = []
    for vp, vp_inv, eye in cams:
        rows.append(np.concatenate([glm.flat(vp), glm.flat(vp_inv), np.asarray(eye, np.float32), [0.0]]))
    return torch.from_numpy(np.asarray(rows, np.float32))


def render(albedo: torch.Tensor, opacity: torch.Tensor, diffuse: torch.Tensor, normal: torch.Tensor,
           depth: torch.Tensor, cameras: torch.Tensor, grid_min, extent: float, params: TraceParams | None = None,
           recompute: bool = False, tape: bool = True) -> torch.Tensor:
    """[V, H, W, 4] = (indirect diffuse rgb, ambient occlusion) per view.

    albedo [R,R,R,3] and opacity [R,R,R] (float32, device) are level 0 of a
    grid whose minimum corner is grid_min and whose edge is extent; the
    G-buffers (diffuse int32 RGBA8 bits, normal int16 half bits, depth float)
    are stacked over V views; cameras is cameras_tensor(...).

    Where the adjoint gets the march's samples: by default, when a gradient
    can follow (grad mode, an input requiring grad), the forward stores them
    (the tape: 16 bytes x cones x (march steps + 2) per pixel) and the
    backward reads them, unless the tape would take more than half of the
    free device memory; with tape=False the backward marches again and keeps
    them per lane; with recompute it fetches them twice (march and reverse
    sweep) and stores nothing."""
    p = params or TraceParams()
    grid = [float(grid_min[0]), float(grid_min[1]), float(grid_min[2]), float(extent)]
    flags = (1 if recompute else 0) | (0 if tape else 2)
    return _vct_torch.render(albedo, opacity, diffuse, normal, depth, cameras, grid, p.as_list(), flags)
