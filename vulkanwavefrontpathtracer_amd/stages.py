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
KEY_ORIGIN_DIR_MORTON60 = 4
KEY_ORIGIN_DIR_MORTON30 = 5


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
        need = _capi.check("wf_sort_scratch_bytes",
                           _capi.load().wf_sort_scratch_bytes(n, kb, key_bits, digit_bits))
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


# ------------------------------------------------------------------ scenes
SCENE_BOXES, SCENE_INTERIOR, SCENE_CITY = 0, 1, 2


class HostScene:
    """Host arrays of a procedural or glTF scene (vertices: [nv, 6] pos+normal)."""

    def __init__(self, kind: int, seed: int, material_groups: int = 0, num_boxes: int = 0,
                 gltf: str | None = None):
        """kind = SCENE_BOXES / SCENE_INTERIOR / SCENE_CITY; num_boxes > 0 builds
        the boxes generator with that many boxes (wf_procedural_boxes);
        gltf = a glTF file to load instead (wf_gltf_load, kind and seed unused)."""
        import numpy as np

        lib = _capi.load()
        nv, nt = C.c_int64(), C.c_int64()
        ng, nm = C.c_int(), C.c_int()
        if gltf is not None:
            gen = lambda *a: call("wf_gltf_load", str(gltf).encode(), *a)  # noqa: E731
        elif num_boxes > 0:
            gen = lambda *a: call("wf_procedural_boxes", seed, num_boxes, *a)  # noqa: E731
        else:
            gen = lambda *a: call("wf_procedural_scene", kind, seed, material_groups, *a)  # noqa: E731
        gen(C.byref(nv), C.byref(nt), C.byref(ng), C.byref(nm),
            None, None, None, None, None, None, None, None)
        self.vertices = np.zeros((nv.value, 6), np.float32)
        self.indices = np.zeros((nt.value, 3), np.uint32)
        self.geom_first = np.zeros(ng.value + 1, np.int32)
        self.tri_material = np.zeros(nt.value, np.int32)
        self.albedo = np.zeros((nm.value, 3), np.float32)
        self.emission = np.zeros((nm.value, 3), np.float32)
        self.bmin = np.zeros(3, np.float32)
        self.bmax = np.zeros(3, np.float32)

        def fp(a):
            return a.ctypes.data_as(_capi.pf)

        gen(None, None, None, None, fp(self.vertices),
             self.indices.ctypes.data_as(_capi.pu32), self.geom_first.ctypes.data_as(_capi.pi32),
             self.tri_material.ctypes.data_as(_capi.pi32), fp(self.albedo), fp(self.emission),
             fp(self.bmin), fp(self.bmax))
        del lib

    @classmethod
    def from_gltf(cls, path) -> "HostScene":
        return cls(0, 0, gltf=str(path))

    @property
    def num_tris(self) -> int:
        return len(self.indices)

    def tri_verts(self):
        """[ntri, 9] float32 vertex positions in original triangle order."""
        return self.vertices[:, :3][self.indices.reshape(-1)].reshape(-1, 9).copy()


class Scene:
    """wf_scene handle (LBVH + triangles + materials) on the device."""

    def __init__(self, vertices, indices, geom_first, tri_material=None, albedo=None,
                 emission=None, device="cuda", vertex_stride=None):
        import numpy as np

        v = torch.as_tensor(np.ascontiguousarray(vertices, np.float32)).to(device)
        idx = torch.as_tensor(np.ascontiguousarray(indices, np.uint32).view(np.int32)).to(device)
        gf = np.ascontiguousarray(geom_first, np.int32)
        mat = None
        if tri_material is not None:
            mat = torch.as_tensor(np.ascontiguousarray(tri_material, np.int32)).to(device)
        stride = vertex_stride or v.shape[1] * 4
        h = C.c_void_p()
        call("wf_scene_create", _ptr(v), v.shape[0], stride, _ptr(idx), idx.numel() // 3,
             gf.ctypes.data_as(_capi.pi32), len(gf) - 1, _ptr(mat), _stream(), C.byref(h))
        self.handle = h
        if albedo is not None:
            a = np.ascontiguousarray(albedo, np.float32)
            e = np.ascontiguousarray(emission, np.float32)
            call("wf_scene_set_materials", h, a.ctypes.data, e.ctypes.data, len(a), _stream())

    @classmethod
    def from_host(cls, hs: "HostScene", device="cuda"):
        return cls(hs.vertices, hs.indices, hs.geom_first, hs.tri_material, hs.albedo,
                   hs.emission, device)

    def bounds(self):
        mn, mx = (C.c_float * 3)(), (C.c_float * 3)()
        call("wf_scene_bounds", self.handle, mn, mx)
        return list(mn), list(mx)

    def info(self):
        nt, nn, nl, ng = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int()
        call("wf_scene_info", self.handle, C.byref(nt), C.byref(nn), C.byref(nl), C.byref(ng))
        return {"tris": nt.value, "nodes": nn.value, "leaves": nl.value, "geoms": ng.value}

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _capi.load().wf_scene_destroy(h)
            except Exception:
                pass
            self.handle = None


def trace(scene: Scene, rays: torch.Tensor, ray_index=None, closest=True, hits=None,
          timed=False, stream=None):
    """Tracer::trace (Tracer/Tracer.cpp:185-207). hits: int32 [n_rays_total, 4]."""
    _check_cuda(rays)
    n = rays.shape[0] if ray_index is None else ray_index.shape[0]
    if hits is None:
        hits = torch.empty((rays.shape[0], 4), dtype=torch.int32, device=rays.device)
    ms = _ms(timed)
    call("wf_trace", scene.handle, _ptr(rays), _ptr(ray_index), _ptr(hits), n,
         1 if closest else 0, _stream(stream), ms)
    return (hits, ms.contents.value) if timed else hits


# ---------------------------------------------------------- ray generation
def init_seeds(n, frame_index, device="cuda"):
    seeds = torch.empty(n, dtype=torch.int32, device=device)
    call("wf_init_seeds", _ptr(seeds), n, frame_index, _stream(), None)
    return seeds


def camera_matrices(position, rotation, fov, aspect, znear, zfar):
    import numpy as np

    v, p, s = (C.c_float * 16)(), (C.c_float * 16)(), (C.c_float * 16)()
    call("wf_camera_matrices", f3(position), f3(rotation), fov, aspect, znear, zfar, v, p, s)
    to = lambda m: np.array(list(m), np.float32).reshape(4, 4).T  # noqa: E731  column-major -> [row, col]
    return to(v), to(p), to(s)


def pixel_table(width, height):
    import numpy as np

    a = np.zeros(width * height, np.int32)
    b = np.zeros(width * height, np.int32)
    call("wf_pixel_table", width, height, a.ctypes.data_as(_capi.pi32), b.ctypes.data_as(_capi.pi32))
    return a, b


def raygen_primary(screen_to_world, origin, index_to_pixel, width, height, sample_index, max_dist):
    """screen_to_world: [4, 4] row/col matrix (as returned by camera_matrices)."""
    import numpy as np

    m = np.ascontiguousarray(np.asarray(screen_to_world, np.float32).T).reshape(-1)
    rays = torch.empty((width * height, 8), dtype=torch.float32, device=index_to_pixel.device)
    call("wf_raygen_primary", m.ctypes.data_as(_capi.pf), f3(origin), _ptr(index_to_pixel),
         _ptr(rays), width, height, sample_index, max_dist, _stream(), None)
    return rays


def raygen_shadow(seeds, in_rays, in_hits, in_s2i, begin, end, num_samples, light, radius):
    n = (end - begin) * num_samples
    dev = in_rays.device
    out = torch.empty((n, 8), dtype=torch.float32, device=dev)
    s2i = torch.empty(n, dtype=torch.int32, device=dev)
    i2s = torch.empty(n, dtype=torch.int32, device=dev)
    call("wf_raygen_shadow", _ptr(seeds), _ptr(in_rays), _ptr(in_hits), _ptr(in_s2i), _ptr(out),
         _ptr(s2i), _ptr(i2s), begin, end, num_samples, f3(light), radius, _stream(), None)
    return out, s2i, i2s


def shade_extend(scene, seeds, in_rays, in_hits, in_s2i, decreases, russian_roulette=False,
                 bounce=0, rr_start=3):
    """RayGen::path (Ray/RayGen.cpp:135-186). Returns (rays, s2i, count)."""
    n = in_rays.shape[0]
    dev = in_rays.device
    out = torch.empty((max(n, 1), 8), dtype=torch.float32, device=dev)
    s2i = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int32, device=dev)
    sb = _capi.load().wf_compact_scratch_bytes(n)
    scratch = torch.empty(max(sb, 256), dtype=torch.uint8, device=dev)
    host = C.c_uint32(0)
    call("wf_shade_extend", scene.handle, _ptr(seeds), _ptr(in_rays), _ptr(in_hits), _ptr(in_s2i),
         n, _ptr(out), _ptr(s2i), _ptr(decreases), _ptr(cnt), 1 if russian_roulette else 0,
         bounce, rr_start, _ptr(scratch), scratch.numel(), C.byref(host), _stream(), None)
    k = host.value
    return out[:k], s2i[:k], k


def fill(buf, value):
    call("wf_fill_f32", _ptr(buf), buf.numel(), float(value), _stream(), None)


def reconstruct_smooth(scene, rays, hits, s2i, pixels, decreases, light, background, spp):
    call("wf_reconstruct_smooth", scene.handle, _ptr(rays), _ptr(hits), _ptr(s2i), _ptr(pixels),
         _ptr(decreases), rays.shape[0], f3(light), f3(background), spp, _stream(), None)


def reconstruct_shadow(shadow_hits, in_s2i, shadow_i2s, in_pixels, out_pixels, begin, end,
                       num_samples, replace):
    call("wf_reconstruct_shadow", _ptr(shadow_hits), _ptr(in_s2i), _ptr(shadow_i2s),
         _ptr(in_pixels), _ptr(out_pixels), begin, end, num_samples, 1 if replace else 0,
         _stream(), None)


def interpolate_colors(frame_pixels, pixels, frame_index, key_value, white_point):
    call("wf_interpolate_colors", _ptr(frame_pixels), _ptr(pixels), pixels.shape[0], frame_index,
         key_value, white_point, _stream(), None)


def count_hits(hits):
    c = torch.zeros(1, dtype=torch.int32, device=hits.device)
    call("wf_count_hits", _ptr(hits), hits.shape[0], _ptr(c), _stream(), None)
    return int(c.item())


def set_tuning(key: str, value: int) -> None:
    """wf_set_tuning: set one process-wide knob (e.g. "sort.onesweep_tiles")."""
    call("wf_set_tuning", key.encode(), int(value))


def get_tuning(key: str) -> int:
    """wf_get_tuning: read one process-wide knob."""
    v = C.c_int()
    call("wf_get_tuning", key.encode(), C.byref(v))
    return v.value
