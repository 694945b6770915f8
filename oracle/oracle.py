"""ctypes view of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the CPU-baseline leg of bench.py, always as the checker.  Arrays are numpy;
rays are float32 [n, 8], hits int32 [n, 4].
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} missing: run `python oracle/build.py`")
        _lib = C.CDLL(str(LIB_PATH))
    return _lib


def _p(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def _rays(rays):
    r = np.ascontiguousarray(rays, dtype=np.float32)
    assert r.ndim == 2 and r.shape[1] == 8
    return r


def keys_aila32(rays, mn, mx):
    r = _rays(rays)
    out = np.empty(len(r), np.uint32)
    load().or_keys_aila32(_p(r), C.c_int64(len(r)), _f3(mn), _f3(mx), _p(out))
    return out


def keys_origin30(rays, mn, mx):
    r = _rays(rays)
    out = np.empty(len(r), np.uint32)
    load().or_keys_origin30(_p(r), C.c_int64(len(r)), _f3(mn), _f3(mx), _p(out))
    return out


def keys_oct60(rays, mn, mx):
    r = _rays(rays)
    out = np.empty(len(r), np.uint64)
    load().or_keys_oct60(_p(r), C.c_int64(len(r)), _f3(mn), _f3(mx), _p(out))
    return out


def keys_morton5d60(rays, mn, mx):
    r = _rays(rays)
    out = np.empty(len(r), np.uint64)
    load().or_keys_morton5d60(_p(r), C.c_int64(len(r)), _f3(mn), _f3(mx), _p(out))
    return out


def keys_morton5d30(rays, mn, mx):
    """WF_KEY_ORIGIN_DIR_MORTON30: the top 30 bits of the 60-bit 5-D Morton
    key (keys.h morton5d30), u32."""
    return (keys_morton5d60(rays, mn, mx) >> np.uint64(30)).astype(np.uint32)


def keys_scene64(rays, scene, mins, maxs):
    r = _rays(rays)
    s = np.ascontiguousarray(scene, np.uint8)
    mn = np.ascontiguousarray(mins, np.float32).reshape(-1)
    mx = np.ascontiguousarray(maxs, np.float32).reshape(-1)
    out = np.empty(len(r), np.uint64)
    load().or_keys_scene64(_p(r), _p(s), C.c_int64(len(r)), _p(mn), _p(mx), _p(out))
    return out


def sort_pairs(keys, vals, key_bits):
    k = np.array(keys, copy=True)
    v = np.array(vals, dtype=np.uint32, copy=True)
    assert k.dtype in (np.uint32, np.uint64, np.int32, np.int64)
    load().or_sort_pairs(_p(k), _p(v), C.c_int64(len(k)), C.c_int(k.itemsize), C.c_int(key_bits))
    return k, v


def reorder(idx, in_rays, in_s2i, i2s_size):
    idx = np.ascontiguousarray(idx, np.uint32)
    r = _rays(in_rays)
    s2i = np.ascontiguousarray(in_s2i, np.int32)
    n = len(idx)
    out_r = np.zeros((n, 8), np.float32)
    out_s2i = np.zeros(n, np.int32)
    out_i2s = np.full(i2s_size, -1, np.int32)
    load().or_reorder(_p(idx), _p(r), _p(out_r), _p(s2i), _p(out_s2i), _p(out_i2s), C.c_int64(n))
    return out_r, out_s2i, out_i2s


# ------------------------------------------------------------------- trace
def trace_f32(rays, tri_verts, closest=True):
    """Brute-force fp32 watertight trace; int32 [n, 4] = {tri, t bits, bary, 0}."""
    r = _rays(rays)
    tv = np.ascontiguousarray(tri_verts, np.float32).reshape(-1, 9)
    out = np.zeros((len(r), 4), np.int32)
    load().or_trace_f32(_p(r), C.c_int64(len(r)), _p(tv), C.c_int64(len(tv)), C.c_int(int(closest)),
                        _p(out))
    return out


def trace_f64(rays, tri_verts):
    """Brute-force fp64 closest hit: (t, tri, edge distance)."""
    r = _rays(rays)
    tv = np.ascontiguousarray(tri_verts, np.float32).reshape(-1, 9)
    t = np.zeros(len(r), np.float64)
    tri = np.zeros(len(r), np.int32)
    edge = np.zeros(len(r), np.float64)
    load().or_trace_f64(_p(r), C.c_int64(len(r)), _p(tv), C.c_int64(len(tv)), _p(t), _p(tri), _p(edge))
    return t, tri, edge


def trace_bvh_f64(rays, tri_verts, closest=True):
    r = _rays(rays)
    tv = np.ascontiguousarray(tri_verts, np.float32).reshape(-1, 9)
    t = np.zeros(len(r), np.float64)
    tri = np.zeros(len(r), np.int32)
    load().or_trace_bvh_f64(_p(r), C.c_int64(len(r)), _p(tv), C.c_int64(len(tv)),
                            C.c_int(int(closest)), _p(t), _p(tri))
    return t, tri


def edge_distance_f64(ray, tri9):
    f = load().or_edge_distance_f64
    f.restype = C.c_double
    r = np.ascontiguousarray(ray, np.float32)
    t = np.ascontiguousarray(tri9, np.float32)
    return f(_p(r), _p(t))


def hits_to_prim_geom(hits_tri, geom_first):
    """Global triangle ids -> (prim in geometry, geometry) as trace.comp reports them."""
    tri = np.asarray(hits_tri)
    g = np.searchsorted(np.asarray(geom_first), tri, side="right") - 1
    prim = tri - np.asarray(geom_first)[np.maximum(g, 0)]
    g = np.where(tri < 0, -1, g)
    prim = np.where(tri < 0, -1, prim)
    return prim.astype(np.int32), g.astype(np.int32)


# ---------------------------------------------------------- ray generation
def tea(a, b, rounds=16):
    f = load().or_tea
    f.restype = C.c_uint32
    return int(f(C.c_uint32(a), C.c_uint32(b), C.c_int(rounds)))


def init_seeds(n, frame):
    out = np.zeros(n, np.uint32)
    load().or_init_seeds(_p(out), C.c_int64(n), C.c_int(frame))
    return out


def halton(i, base):
    f = load().or_halton
    f.restype = C.c_float
    return float(f(C.c_int(i), C.c_int(base)))


def sobol2(i):
    xy = np.zeros(2, np.float32)
    load().or_sobol2(C.c_uint32(i), _p(xy))
    return xy


def raygen_primary(s2w_colmajor, origin, i2p, w, h, sample_index, max_dist):
    m = np.ascontiguousarray(s2w_colmajor, np.float32).reshape(16)
    o = np.ascontiguousarray(origin, np.float32)
    lut = np.ascontiguousarray(i2p, np.int32)
    out = np.zeros((w * h, 8), np.float32)
    load().or_raygen_primary(_p(m), _p(o), _p(lut), C.c_int(w), C.c_int(h), C.c_int(sample_index),
                             C.c_float(max_dist), _p(out))
    return out


def raygen_shadow(seeds, in_rays, in_hits, in_s2i, begin, end, S, light, radius):
    n = (end - begin) * S
    out = np.zeros((n, 8), np.float32)
    s2i = np.zeros(n, np.int32)
    i2s = np.zeros(n, np.int32)
    sd = np.ascontiguousarray(seeds, np.uint32)
    r = _rays(in_rays)
    h = np.ascontiguousarray(in_hits, np.int32)
    s = np.ascontiguousarray(in_s2i, np.int32)
    lt = np.ascontiguousarray(light, np.float32)
    load().or_raygen_shadow(_p(sd), _p(r), _p(h), _p(s), C.c_int64(begin), C.c_int64(end), C.c_int(S),
                            _p(lt), C.c_float(radius), _p(out), _p(s2i), _p(i2s))
    return out, s2i, i2s


def tri_normals(vertices, indices):
    v = np.asarray(vertices, np.float32)
    return v[:, 3:6][np.asarray(indices).reshape(-1)].reshape(-1, 9).copy()


def shade_extend(geom_first, tri_n, tri_mat, albedo, seeds, in_rays, in_hits, in_s2i, decreases, rr):
    gf = np.ascontiguousarray(geom_first, np.int32)
    tn = np.ascontiguousarray(tri_n, np.float32)
    tm = np.ascontiguousarray(tri_mat, np.int32)
    al = np.ascontiguousarray(albedo, np.float32)
    sd = np.array(seeds, dtype=np.uint32, copy=True)
    r = _rays(in_rays)
    h = np.ascontiguousarray(in_hits, np.int32)
    s = np.ascontiguousarray(in_s2i, np.int32)
    dec = np.array(decreases, dtype=np.float32, copy=True)
    n = len(r)
    out = np.zeros((max(n, 1), 8), np.float32)
    s2i = np.zeros(max(n, 1), np.int32)
    f = load().or_shade_extend
    f.restype = C.c_int64
    k = f(_p(gf), _p(tn), _p(tm), _p(al), C.c_int(len(al)), _p(sd), _p(r), _p(h), _p(s), C.c_int64(n),
          _p(out), _p(s2i), _p(dec), C.c_int(int(rr)))
    return out[:k], s2i[:k], sd, dec


def reconstruct_smooth(geom_first, tri_n, tri_mat, emission, rays, hits, s2i, pixels, decreases,
                       light, bg, spp):
    px = np.array(pixels, dtype=np.float32, copy=True)
    args = [np.ascontiguousarray(a, dt) for a, dt in
            [(geom_first, np.int32), (tri_n, np.float32), (tri_mat, np.int32), (emission, np.float32)]]
    r = _rays(rays)
    h = np.ascontiguousarray(hits, np.int32)
    s = np.ascontiguousarray(s2i, np.int32)
    dec = np.ascontiguousarray(decreases, np.float32)
    lt = np.ascontiguousarray(light, np.float32)
    b = np.ascontiguousarray(bg, np.float32)
    load().or_reconstruct_smooth(_p(args[0]), _p(args[1]), _p(args[2]), _p(args[3]),
                                 C.c_int(len(args[3]) // 3 if args[3].ndim == 1 else len(args[3])),
                                 _p(r), _p(h), _p(s), _p(px), _p(dec), C.c_int64(len(r)), _p(lt), _p(b),
                                 C.c_int(spp))
    return px


def reconstruct_shadow(shadow_hits, in_s2i, shadow_i2s, in_px, out_px, begin, end, S, replace):
    o = np.array(out_px, dtype=np.float32, copy=True)
    sh = np.ascontiguousarray(shadow_hits, np.int32)
    s = np.ascontiguousarray(in_s2i, np.int32)
    i2s = np.ascontiguousarray(shadow_i2s, np.int32)
    ip = np.ascontiguousarray(in_px, np.float32)
    load().or_reconstruct_shadow(_p(sh), _p(s), _p(i2s), _p(ip), _p(o), C.c_int64(begin),
                                 C.c_int64(end), C.c_int(S), C.c_int(int(replace)))
    return o


def tonemap(rgb, key, white):
    i = np.ascontiguousarray(rgb, np.float32)
    o = np.zeros(3, np.float32)
    load().or_tonemap(_p(i), C.c_float(key), C.c_float(white), _p(o))
    return o


def interpolate(frame, pixels, frame_index, key, white):
    f = np.array(frame, dtype=np.float32, copy=True)
    p = np.array(pixels, dtype=np.float32, copy=True)
    load().or_interpolate(_p(f), _p(p), C.c_int64(len(f)), C.c_int(frame_index), C.c_float(key),
                          C.c_float(white))
    return f, p


class FrameCfg(C.Structure):
    _fields_ = [("w", C.c_int), ("h", C.c_int), ("spp", C.c_int), ("depth", C.c_int), ("S", C.c_int),
                ("frame", C.c_int), ("rr", C.c_int), ("rr_start", C.c_int),
                ("light", C.c_float * 3), ("bg", C.c_float * 3), ("origin", C.c_float * 3),
                ("radius", C.c_float), ("max_dist", C.c_float), ("s2w", C.c_float * 16),
                ("pixel_stride", C.c_int)]


def render_frame_f64(tri_verts, tri_n, tri_mat, albedo, emission, *, w, h, spp, depth, S, frame,
                     light, bg, origin, radius, max_dist, s2w_colmajor, rr=False, rr_start=3,
                     d_pixels=None, hints=None, first_disagreement=None, pixel_stride=1):
    """Frame 1 of Renderer::renderPath, primary rays in fp32 (as raygenPrimary
    makes them) and the paths followed in fp64.  Returns ([w*h, 4] float32
    linear RGBA, [w*h] bool ambiguous): a pixel is ambiguous when some vertex
    of its paths grazes a triangle edge, ties two hits or sits on an interval
    end within the fp32 error bound propagated along the path, where an fp32
    pipeline may follow another path.

    hints = (hit_tri [spp, depth+1, w*h] int32, shadow_visible [spp, depth+1,
    w*h, S] uint8[, bounce_signs [spp, depth+1, w*h] uint8]), the fp32
    pipeline's discrete decisions (the renderer's decision log): where one
    differs from the fp64 decision the replay takes it, and a third array
    [w*h] uint8 is returned with bit 0 set where that happened at an
    ambiguous decision, bit 1 where it happened at an unambiguous one (a real
    disagreement), bit 2 where a bounce sign decision (faceforward, basis
    branch) lies within the fp32 bound and no bounce_signs were given, and
    bit 3 where the path's propagated fp32 uncertainty passes 1e-3 world
    units at some vertex (beyond it the two paths part continuously).  first_disagreement (optional
    [w*h, 4] int32 array) receives, per pixel, the first unambiguous
    disagreement as (vertex, kind: 1 hit / 2+i shadow sample i / -1
    roulette, fp64 choice, fp32 choice), vertex -1 where there is none.

    With d_pixels ([w*h, 4] float32, dLoss/dpixel), also returns the fp64
    adjoint: (image, ambiguous, grad_albedo [nmat, 3], grad_emission [nmat, 3])
    with grad = sum over pixels of d_pixels . dpixel/dparameter.

    pixel_stride k > 1 renders only pixels 0, k, 2k, ... (a sample of a
    large frame); the other entries of every output stay as initialised."""
    cfg = FrameCfg(w, h, spp, depth, S, frame, int(rr), rr_start, (C.c_float * 3)(*light),
                   (C.c_float * 3)(*bg), (C.c_float * 3)(*origin), radius, max_dist,
                   (C.c_float * 16)(*np.asarray(s2w_colmajor, np.float32).reshape(16)),
                   int(pixel_stride))
    tv = np.ascontiguousarray(tri_verts, np.float32)
    tn = np.ascontiguousarray(tri_n, np.float32)
    tm = np.ascontiguousarray(tri_mat, np.int32)
    al = np.ascontiguousarray(albedo, np.float32)
    em = np.ascontiguousarray(emission, np.float32)
    out = np.zeros((w * h, 4), np.float32)
    amb = np.zeros(w * h, np.uint8)
    ht = hv = hs = fl = None
    if hints is not None:
        ht = np.ascontiguousarray(hints[0], np.int32).reshape(spp, depth + 1, w * h)
        hv = np.ascontiguousarray(hints[1], np.uint8).reshape(spp, depth + 1, w * h, S)
        if len(hints) > 2 and hints[2] is not None:
            hs = np.ascontiguousarray(hints[2], np.uint8).reshape(spp, depth + 1, w * h)
        fl = np.zeros(w * h, np.uint8)
    fd = first_disagreement
    pp = lambda a: _p(a) if a is not None else None  # noqa: E731
    hint_args = (pp(ht), pp(hv), pp(hs), pp(fl), pp(fd))
    if d_pixels is None:
        load().or_render_frame_f64(_p(tv), _p(tn), _p(tm), C.c_int64(len(tv)), _p(al), _p(em),
                                   C.c_int(len(al)), C.byref(cfg), _p(out), _p(amb), None, None, None,
                                   *hint_args)
        return (out, amb.astype(bool)) if fl is None else (out, amb.astype(bool), fl)
    dp = np.ascontiguousarray(d_pixels, np.float32).reshape(w * h, 4)
    ga = np.zeros((len(al), 3), np.float64)
    ge = np.zeros((len(al), 3), np.float64)
    load().or_render_frame_f64(_p(tv), _p(tn), _p(tm), C.c_int64(len(tv)), _p(al), _p(em),
                               C.c_int(len(al)), C.byref(cfg), _p(out), _p(amb), _p(dp), _p(ga), _p(ge),
                               *hint_args)
    return out, amb.astype(bool), ga, ge


def trace_bvh_f32(rays, tri_verts, closest=True):
    """fp32 watertight trace over a BVH (large scenes); same records as
    trace_f32."""
    r = _rays(rays)
    tv = np.ascontiguousarray(tri_verts, np.float32).reshape(-1, 9)
    out = np.zeros((len(r), 4), np.int32)
    load().or_trace_bvh_f32(_p(r), C.c_int64(len(r)), _p(tv), C.c_int64(len(tv)),
                            C.c_int(int(closest)), _p(out))
    return out


class Bvh:
    """The fp32 trace's BVH built once (or_bvh_create) for repeated traces."""

    def __init__(self, tri_verts):
        self.tv = np.ascontiguousarray(tri_verts, np.float32).reshape(-1, 9)
        lib = load()
        lib.or_bvh_create.restype = C.c_void_p
        self.h = C.c_void_p(lib.or_bvh_create(_p(self.tv), C.c_int64(len(self.tv))))

    def trace_f32(self, rays, closest=True):
        r = _rays(rays)
        out = np.zeros((len(r), 4), np.int32)
        load().or_bvh_trace_f32(self.h, _p(r), C.c_int64(len(r)), C.c_int(int(closest)), _p(out))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            load().or_bvh_destroy(self.h)
            self.h = None


def lcg(seed, n):
    """n draws of raygenPath's LCG from seed: (states uint32, floats)."""
    st = np.zeros(n, np.uint32)
    r = np.zeros(n, np.float32)
    load().or_lcg(C.c_uint32(seed), C.c_int64(n), _p(st), _p(r))
    return st, r


def duff_basis(n, f64=False):
    """(b1, b2) of the Duff et al. basis around unit normal n."""
    dt = np.float64 if f64 else np.float32
    nn = np.ascontiguousarray(n, dt)
    b1 = np.zeros(3, dt)
    b2 = np.zeros(3, dt)
    (load().or_duff_basis_f64 if f64 else load().or_duff_basis)(_p(nn), _p(b1), _p(b2))
    return b1, b2


def cosine_dir(r0, r1, n, f64=False):
    """raygenPath's cosine-weighted bounce direction about unit normal n."""
    dt = np.float64 if f64 else np.float32
    nn = np.ascontiguousarray(n, dt)
    out = np.zeros(3, dt)
    if f64:
        load().or_cosine_dir_f64(C.c_double(r0), C.c_double(r1), _p(nn), _p(out))
    else:
        load().or_cosine_dir(C.c_float(r0), C.c_float(r1), _p(nn), _p(out))
    return out


def shadow_point(seed, pixel, i, S):
    """Shadow sample i of S of a pixel in [-1,1]^3 (Sobol x/y, Hammersley z,
    Cranley-Patterson offset from the Jenkins-mixed pixel seed)."""
    p = np.zeros(3, np.float32)
    load().or_shadow_point(C.c_uint32(seed), C.c_int(pixel), C.c_int(i), C.c_int(S), _p(p))
    return p


def jenkins(a, b, c):
    abc = np.array([a, b, c], np.uint32)
    load().or_jenkins(_p(abc))
    return tuple(int(x) for x in abc)
