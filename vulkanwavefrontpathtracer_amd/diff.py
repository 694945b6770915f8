"""Differentiable rendering: torch autograd over the frame loop.

The autograd Function is libtorch C++ (torch_ext/render_autograd.cpp, built
into _wfpt_torch.so): its forward pass is frame 1 of the C++ wavefront frame
loop (wf_renderer_render) and its backward pass the path-replay adjoint
kernel (wf_renderer_adjoint), which recomputes every path vertex under the
frame's random numbers and returns dLoss/dalbedo and dLoss/demission per
material.  Parameters are [num_materials, 3] float32 GPU tensors."""
from __future__ import annotations

import torch

from .renderer import Renderer


def _ext():
    try:
        from . import _wfpt_torch
    except ImportError as e:  # built by vulkanwavefrontpathtracer_amd.build.build_torch_ext
        raise ImportError("_wfpt_torch.so is not built: run __graft_entry__.build()") from e
    return _wfpt_torch


def render(renderer: Renderer, albedo: torch.Tensor, emission: torch.Tensor) -> torch.Tensor:
    """image[h, w, 3] (linear frame estimate) under the renderer's env settings."""
    return _ext().render(albedo, emission, renderer.handle.value)
