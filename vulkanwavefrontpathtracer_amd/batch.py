"""Python handle on the batched wavefront queue (wf_batch_*): many scenes and
views of uneven size rendered in one persistent queue, with a scene-ID
prefix in the ray sort key, a HIP-graph-captured pass and an autotuner over
sort digit width, trace block size and reorder period."""
from __future__ import annotations

import ctypes as C
import json
import math

import numpy as np

from . import _capi
from ._capi import call


class WfView(C.Structure):
    _fields_ = [("scene", C.c_int), ("width", C.c_int), ("height", C.c_int), ("spp", C.c_int),
                ("position", C.c_float * 3), ("rotation", C.c_float * 3), ("fov", C.c_float),
                ("znear", C.c_float), ("zfar", C.c_float), ("light", C.c_float * 3),
                ("background", C.c_float * 3), ("shadow_radius", C.c_float)]


def make_view(scene, width, height, spp, position, rotation=(0, 0, 0), fov=60.0, znear=0.1,
              zfar=512.0, light=(0, 10.4, 0), background=(0.529, 0.808, 0.98),
              shadow_radius=1.5) -> WfView:
    f3 = lambda v: (C.c_float * 3)(*[float(x) for x in v])  # noqa: E731
    return WfView(scene, width, height, spp, f3(position), f3(rotation), fov, znear, zfar,
                  f3(light), f3(background), shadow_radius)


class Batch:
    def __init__(self, scenes, views, depth=6, shadow_samples=2):
        self._scenes = list(scenes)          # keep the scenes alive
        self.views = list(views)
        arr = (C.c_void_p * len(self._scenes))(*[s.handle.value for s in self._scenes])
        varr = (WfView * len(self.views))(*self.views)
        h = C.c_void_p()
        call("wf_batch_create", arr, len(self._scenes), C.cast(varr, C.c_void_p), len(self.views),
             depth, shadow_samples, C.byref(h))
        self.handle = h

    def set_tuning(self, digit_bits=8, block=64, reorder_period=1):
        call("wf_batch_set_tuning", self.handle, digit_bits, block, reorder_period)

    def tuning(self):
        d, b, p = C.c_int(), C.c_int(), C.c_int()
        call("wf_batch_get_tuning", self.handle, C.byref(d), C.byref(b), C.byref(p))
        return {"digit_bits": d.value, "block": b.value, "reorder_period": p.value}

    def autotune(self) -> float:
        ms = C.c_float()
        call("wf_batch_autotune", self.handle, C.byref(ms))
        return ms.value

    def set_graph(self, on=True):
        call("wf_batch_set_graph", self.handle, 1 if on else 0)

    def render(self) -> float:
        ms = C.c_float()
        call("wf_batch_render", self.handle, C.byref(ms))
        return ms.value

    def pixels(self, view) -> np.ndarray:
        v = self.views[view]
        out = np.zeros((v.width * v.height, 4), np.float32)
        call("wf_batch_read_pixels", self.handle, view, out.ctypes.data_as(_capi.pf))
        return out

    def stats(self):
        c = (C.c_uint64 * 3)()
        tp = C.c_int64()
        call("wf_batch_stats", self.handle, c, C.byref(tp))
        return {"primary_rays": c[0], "shadow_rays": c[1], "path_rays": c[2],
                "pixels": tp.value}

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _capi.load().wf_batch_destroy(h)
            except Exception:
                pass
            self.handle = None


def workload(seed=3, num_views=64, tri_range=(1_000, 500_000),
             resolutions=(128, 256, 512, 1024), spp_max=16):
    """The batched workload: num_views scenes from the boxes generator with
    triangle counts log-uniform in tri_range (scene k seeded 3000 + k), each
    seen by one view with a resolution drawn from `resolutions` (square) and
    spp uniform in 1..spp_max; camera, light and background of the boxes
    scene's env.  Returns (box counts, scene seeds, view dicts)."""
    rng = np.random.default_rng(seed)
    lo, hi = math.log(tri_range[0]), math.log(tri_range[1])
    boxes, seeds, views = [], [], []
    for k in range(num_views):
        tris = math.exp(rng.uniform(lo, hi))
        boxes.append(max(1, int(round((tris - 4) / 192))))
        seeds.append(3000 + k)
        res = int(rng.choice(resolutions))
        views.append(dict(scene=k, width=res, height=res, spp=int(rng.integers(1, spp_max + 1))))
    return boxes, seeds, views


def boxes_env_view(env_path, scene, width, height, spp):
    """A view with the camera, light, background and shadow radius of a
    boxes-scene env file."""
    cfg = json.loads(open(env_path).read())
    vec = lambda s: [float(x) for x in s.split()]  # noqa: E731
    cam, sc, R = cfg["Camera"], cfg["Scene"], cfg["Renderer"]
    return make_view(scene, width, height, spp, vec(cam["position"]), vec(cam["rotation"]),
                     cam["fieldOfView"], cam["nearPlane"], cam["farPlane"], vec(sc["light"]),
                     vec(sc["backgroundcolor"]), R["shadowRadius"])
