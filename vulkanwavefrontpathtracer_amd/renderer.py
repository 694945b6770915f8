"""Python handle on the C++ frame loop (wf_renderer_*), mirroring Renderer
(projects/WaveFrontPathTracer/Renderer/Renderer.h:212-276) and the
Benchmark report fields (Benchmark/Benchmark.cpp:128-195)."""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import numpy as np

from . import _capi
from ._capi import call

ENV_DIR = Path(__file__).resolve().parent / "env"


def env_path(name: str) -> Path:
    """Path of a bundled env file (config0.json ... config4.json)."""
    p = ENV_DIR / name
    if not p.exists():
        raise FileNotFoundError(p)
    return p


class _SortLog(C.Structure):
    _fields_ = [("ray_count", C.c_int64), ("sorted_rays", C.c_int64), ("morton_ms", C.c_float),
                ("sort_ms", C.c_float), ("reorder_ms", C.c_float), ("trace_sort_ms", C.c_float),
                ("trace_ms", C.c_float), ("reserved", C.c_float)]


class Renderer:
    def __init__(self, env: str | Path | dict, overrides: dict | None = None,
                 asset_path: str | Path | None = None):
        """env: an env.json path or its parsed dict; overrides: dotted keys
        ("Scene.filename": ...) set on top; asset_path: prefix of a relative
        Scene.filename (default: the env file's ../assets/)."""
        h = C.c_void_p()
        if isinstance(env, dict) or overrides:
            data = env if isinstance(env, dict) else json.loads(Path(env).read_text())
            data = json.loads(json.dumps(data))
            for dotted, value in (overrides or {}).items():
                node = data
                *path, leaf = dotted.split(".")
                for k in path:
                    node = node.setdefault(k, {})
                node[leaf] = value
            call("wf_renderer_create_from_string", json.dumps(data).encode(), C.byref(h))
        else:
            call("wf_renderer_create", str(env).encode(), C.byref(h))
        self.handle = h
        if asset_path is None and not isinstance(env, dict):
            asset_path = Path(env).resolve().parent.parent / "assets"
        if asset_path is not None:
            call("wf_renderer_set_asset_path", self.handle, str(asset_path).encode())

    def build_scene(self):
        call("wf_renderer_build_scene", self.handle)
        return self

    def set_scene(self, scene, scene_min=None, scene_max=None):
        mn = _capi.f3(scene_min) if scene_min is not None else None
        mx = _capi.f3(scene_max) if scene_max is not None else None
        call("wf_renderer_set_scene", self.handle, scene.handle, mn, mx)
        self._scene_ref = scene

    def render(self) -> float:
        ms = C.c_float(0)
        call("wf_renderer_render", self.handle, C.byref(ms))
        return ms.value

    def reset_frame_index(self):
        call("wf_renderer_reset_frame_index", self.handle)

    @property
    def size(self):
        w, h = C.c_int(), C.c_int()
        call("wf_renderer_size", self.handle, C.byref(w), C.byref(h))
        return w.value, h.value

    def pixels(self, tone_mapped=False) -> np.ndarray:
        w, h = self.size
        out = np.zeros((h * w, 4), np.float32)
        call("wf_renderer_read_pixels", self.handle, 1 if tone_mapped else 0,
             out.ctypes.data_as(_capi.pf))
        return out

    def pixels_tensor(self, tone_mapped=False):
        """The framebuffer as a [h, w, 4] float32 torch tensor on the GPU (a
        copy; the renderer's buffer is reused by the next frame)."""
        import torch

        w, h = self.size
        p = C.c_void_p()
        call("wf_renderer_pixels_device", self.handle, 1 if tone_mapped else 0, C.byref(p))
        out = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
        _capi.memcpy_d2d(out.data_ptr(), p.value, out.numel() * 4)
        return out

    def set_materials(self, albedo, emission):
        """Per-material albedo and emission ([nmat, 3] float32, host arrays or
        GPU tensors) of the renderer's scene."""
        import torch

        a = albedo.detach().float().contiguous() if torch.is_tensor(albedo) else \
            np.ascontiguousarray(albedo, np.float32)
        e = emission.detach().float().contiguous() if torch.is_tensor(emission) else \
            np.ascontiguousarray(emission, np.float32)
        pa = a.data_ptr() if torch.is_tensor(a) else a.ctypes.data
        pe = e.data_ptr() if torch.is_tensor(e) else e.ctypes.data
        if torch.is_tensor(a) and a.is_cuda:
            torch.cuda.synchronize()
        call("wf_scene_set_materials", self.scene_handle(), pa, pe, a.shape[0], None)

    def adjoint(self, d_pixels, replay=False, frame=0):
        """Path-replay adjoint of the last frame: d_pixels = dLoss/dpixel of
        the linear frame estimate ([h, w, 4] or [h*w, 4] float32 on the GPU).
        Returns (grad_albedo, grad_emission[, replayed frame]) as tensors, and
        the kernel's GPU milliseconds in self.last_adjoint_ms."""
        import torch

        nm = self.num_materials()
        dp = d_pixels.detach().float().contiguous()
        ga = torch.empty((nm, 3), dtype=torch.float32, device=dp.device)
        ge = torch.empty_like(ga)
        rp = torch.empty_like(dp) if replay else None
        ms = C.c_float(0)
        torch.cuda.synchronize()
        call("wf_renderer_adjoint", self.handle, frame, dp.data_ptr(), ga.data_ptr(), ge.data_ptr(),
             rp.data_ptr() if replay else None, C.byref(ms))
        self.last_adjoint_ms = ms.value
        return (ga, ge, rp) if replay else (ga, ge)

    def num_materials(self):
        n = C.c_int()
        call("wf_scene_get_materials", self.scene_handle(), C.byref(n), None, None)
        return n.value

    def materials(self):
        """(albedo, emission) of the renderer's scene as [nmat, 3] numpy arrays."""
        nm = self.num_materials()
        a = np.zeros((nm, 3), np.float32)
        e = np.zeros((nm, 3), np.float32)
        n = C.c_int()
        call("wf_scene_get_materials", self.scene_handle(), C.byref(n), a.ctypes.data, e.ctypes.data)
        return a, e

    def save_image(self, path, linear=False):
        """Tone-mapped PNG (present.comp's rgba8 conversion) or linear PFM."""
        call("wf_renderer_save_image", self.handle, str(path).encode(), 1 if linear else 0)

    def stats(self) -> dict:
        counts = (C.c_uint64 * 3)()
        times = (C.c_float * 3)()
        stage = (C.c_float * 8)()
        sorted_rays = C.c_uint64()
        call("wf_renderer_stats", self.handle, counts, times, stage, C.byref(sorted_rays))
        names = ["keys", "sort", "reorder", "trace_sorted", "trace_unsorted", "shade", "raygen",
                 "reconstruct"]
        return {
            "primary_rays": counts[0], "shadow_rays": counts[1], "path_rays": counts[2],
            "primary_ms": times[0], "shadow_ms": times[1], "path_ms": times[2],
            "stage_ms": dict(zip(names, list(stage))), "sorted_path_rays": sorted_rays.value,
        }

    def env_int(self, name):
        v = C.c_int()
        call("wf_env_get_int", self.handle, name.encode(), C.byref(v))
        return v.value

    def env_float(self, name):
        v = C.c_float()
        call("wf_env_get_float", self.handle, name.encode(), C.byref(v))
        return v.value

    def env_string(self, name):
        buf = C.create_string_buffer(512)
        call("wf_env_get_string", self.handle, name.encode(), buf, 512)
        return buf.value.decode()

    def env_vector(self, name):
        v = (C.c_float * 3)()
        call("wf_env_get_vector", self.handle, name.encode(), v)
        return list(v)

    def sort_logs(self, kind="path") -> list[dict]:
        """SortLog entries of the last frame (wf_renderer_sort_logs): per path
        bounce, or per shadow vertex (kind="shadow")."""
        k = 0 if kind == "path" else 1
        n = C.c_int()
        call("wf_renderer_sort_logs", self.handle, k, None, 0, C.byref(n))
        arr = (_SortLog * max(n.value, 1))()
        call("wf_renderer_sort_logs", self.handle, k, C.cast(arr, C.c_void_p), n.value, None)
        return [{f: getattr(arr[i], f) for f, _ in _SortLog._fields_ if f != "reserved"}
                for i in range(n.value)]

    def set_decision_log(self, on=True):
        call("wf_renderer_set_decision_log", self.handle, 1 if on else 0)

    def decisions(self):
        """(hit_tri [spp, depth+1, h*w] int32, shadow_visible [spp, depth+1, h*w, S] uint8,
        bounce_signs [spp, depth+1, h*w] uint8) of the last frame rendered with the
        decision log on."""
        spp = self.env_int("Renderer.numberOfPrimarySamples")
        depth = self.env_int("Renderer.recursionDepth")
        S = self.env_int("Renderer.numberOfShadowSamples")
        w, h = self.size
        tri = np.zeros((spp, depth + 1, w * h), np.int32)
        vis = np.zeros((spp, depth + 1, w * h, S), np.uint8)
        sign = np.zeros((spp, depth + 1, w * h), np.uint8)
        call("wf_renderer_decisions", self.handle, tri.ctypes.data, vis.ctypes.data, sign.ctypes.data)
        return tri, vis, sign

    def set_capture(self, on=True):
        call("wf_renderer_set_capture", self.handle, 1 if on else 0)

    def captured(self, bounce):
        """(rays device address, slot-to-pixel device address, count) of a bounce."""
        p, q, n = C.c_void_p(), C.c_void_p(), C.c_int64()
        call("wf_renderer_captured", self.handle, bounce, C.byref(p), C.byref(q), C.byref(n))
        return p.value, q.value, n.value

    def scene_handle(self):
        p = C.c_void_p()
        call("wf_renderer_get_scene", self.handle, C.byref(p))
        return p

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _capi.load().wf_renderer_destroy(h)
            except Exception:
                pass
            self.handle = None
