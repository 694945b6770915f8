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


class Renderer:
    def __init__(self, env: str | Path | dict, overrides: dict | None = None):
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
