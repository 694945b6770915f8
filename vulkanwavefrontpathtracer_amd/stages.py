"""Python view of the stage API (Tracer / RayGen / reconstruction) over torch
device tensors.  Every function calls the HIP implementation in libwfpt.so
through the C ABI; PyTorch only provides device memory and the stream.

Rays are float32 tensors of shape [n, 8] (Ray.h:32-37 layout), hits are int32
tensors of shape [n, 4] (Ray.h:39-44 layout; column 2 holds the float bits of
t, read it with hits_t()).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _capi
from ._capi import call, f3

KEY_AILA32 = 0
KEY_ORIGIN30 = 1
KEY_ORIGIN30_OCT30 = 2
KEY_SCENE6_ORIGIN30_OCT28 = 3


def _ptr(t: torch.Tensor | None) -> C.c_void_p:
    if t is None:
        return C.c_void_p(0)
    return C.c_void_p(t.data_ptr())


def _stream(stream=None) -> C.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _check_cuda(*ts):
    for t in ts:
        if t is not None and not t.is_cuda:
            raise ValueError("device tensor expected")


def _ms(timed: bool):
    return C.pointer(C.c_float(0.0)) if timed else None


def key_dtype(mode: int) -> torch.dtype:
    nbytes = _capi.load().wf_key_bytes(mode)
    if nbytes < 0:
        raise ValueError(f"unknown key mode {mode}")
    return torch.int32 if nbytes == 4 else torch.int64


def key_bits(mode: int) -> int:
    return int(call("wf_key_bits", mode))


def morton_keys(rays: torch.Tensor, mode: int, scene_min, scene_max, keys=None, indices=None,
                timed: bool = False, stream=None):
    """Tracer::computeMortonCodes (Tracer/Tracer.cpp:22-37). Returns (keys, indices[, ms])."""
    _check_cuda(rays)
    n = rays.shape[0]
    if keys is None:
        keys = torch.empty(n, dtype=key_dtype(mode), device=rays.device)
    if indices is None:
        indices = torch.empty(n, dtype=torch.int32, device=rays.device)
    ms = _ms(timed)
    call("wf_morton_keys", _ptr(rays), _ptr(keys), _ptr(indices), n, mode, f3(scene_min),
         f3(scene_max), _stream(stream), ms)
    return (keys, indices, ms.contents.value) if timed else (keys, indices)


def morton_keys_batched(rays, scene_of_ray, scene_mins, scene_maxs, timed=False, stream=None):
    _check_cuda(rays, scene_of_ray)
    n = rays.shape[0]
    keys = torch.empty(n, dtype=torch.int64, device=rays.device)
    idx = torch.empty(n, dtype=torch.int32, device=rays.device)
    ns = len(scene_mins)
    mn = (C.c_float * (3 * ns))(*[float(x) for v in scene_mins for x in v])
    mx = (C.c_float * (3 * ns))(*[float(x) for v in scene_maxs for x in v])
    ms = _ms(timed)
    call("wf_morton_keys_batched", _ptr(rays), _ptr(scene_of_ray), _ptr(keys), _ptr(idx), n,
         mn, mx, ns, _stream(stream), ms)
    return (keys, idx, ms.contents.value) if timed else (keys, idx)


class Sorter:
    """Owns the scratch of wf_sort_pairs (VrdxSorter + storage buffer)."""

    def __init__(self):
        self.scratch = None

    def sort(self, keys: torch.Tensor, vals: torch.Tensor, key_bits: int, digit_bits: int = 8,
             timed: bool = False, stream=None):
        _check_cuda(keys, vals)
        n = keys.shape[0]
        kb = keys.element_size()
        need = _capi.load().wf_sort_scratch_bytes(n, kb, key_bits, digit_bits)
        if need < 0:
            raise ValueError(_capi.load().wf_last_error().decode())
        if self.scratch is None or self.scratch.numel() < need:
            self.scratch = torch.empty(max(need, 256), dtype=torch.uint8, device=keys.device)
        ms = _ms(timed)
        call("wf_sort_pairs", _ptr(keys), _ptr(vals), n, kb, key_bits, digit_bits,
             _ptr(self.scratch), self.scratch.numel(), _stream(stream), ms)
        return ms.contents.value if timed else None


def sort_pairs(keys, vals, key_bits, digit_bits=8, timed=False):
    return Sorter().sort(keys, vals, key_bits, digit_bits, timed)


def reorder(idx, in_rays, in_s2i, i2s_size=None, out_rays=None, out_s2i=None, out_i2s=None,
            timed=False, stream=None):
    """Tracer::reorderRays (Tracer/Tracer.cpp:80-98). Returns (rays, s2i, i2s[, ms])."""
    _check_cuda(idx, in_rays, in_s2i)
    n = idx.shape[0]
    dev = in_rays.device
    if out_rays is None:
        out_rays = torch.empty_like(in_rays[:n])
    if out_s2i is None:
        out_s2i = torch.empty(n, dtype=torch.int32, device=dev)
    if out_i2s is None:
        size = i2s_size if i2s_size is not None else n
        out_i2s = torch.full((size,), -1, dtype=torch.int32, device=dev)
    ms = _ms(timed)
    call("wf_reorder", _ptr(idx), _ptr(in_rays), _ptr(out_rays), _ptr(in_s2i), _ptr(out_s2i),
         _ptr(out_i2s), n, _stream(stream), ms)
    res = (out_rays, out_s2i, out_i2s)
    return res + (ms.contents.value,) if timed else res


def hits_t(hits: torch.Tensor) -> torch.Tensor:
    return hits[:, 2].view(torch.float32)
