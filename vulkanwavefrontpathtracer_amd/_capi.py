"""ctypes bindings of libwfpt.so (the C ABI declared in include/wfpt.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libwfpt.so"

vp = C.c_void_p
i32 = C.c_int
i64 = C.c_int64
u64 = C.c_uint64
f32 = C.c_float
pf = C.POINTER(C.c_float)
pi32 = C.POINTER(C.c_int32)
pu32 = C.POINTER(C.c_uint32)
pi64 = C.POINTER(C.c_int64)
pu64 = C.POINTER(C.c_uint64)
cs = C.c_char_p


class WfRay(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("ox", "oy", "oz", "tmin", "dx", "dy", "dz", "tmax")]


class WfRayBuffer(C.Structure):
    _fields_ = [
        ("rays", vp * 2),
        ("index_to_slot", vp * 2),
        ("slot_to_index", vp * 2),
        ("results", vp),
        ("keys", vp),
        ("indices", vp),
        ("sort_scratch", vp),
        ("sort_scratch_bytes", i64),
        ("size", i64),
        ("capacity", i64),
        ("swap", C.c_int32),
        ("closest_hit", C.c_int32),
    ]


# name -> (restype, argtypes); every symbol here is declared in include/wfpt.h.
SIGNATURES: dict[str, tuple] = {
    "wf_status_string": (cs, [i32]),
    "wf_last_error": (cs, []),
    "wf_version": (cs, []),
    "wf_key_bytes": (i32, [i32]),
    "wf_key_bits": (i32, [i32]),
    "wf_morton_keys": (i32, [vp, vp, vp, i64, i32, pf, pf, vp, pf]),
    "wf_morton_keys_batched": (i32, [vp, vp, vp, vp, i64, pf, pf, i32, vp, pf]),
    "wf_sort_scratch_bytes": (i64, [i64, i32, i32, i32]),
    "wf_sort_pairs": (i32, [vp, vp, i64, i32, i32, i32, vp, i64, vp, pf]),
    "wf_sort_pairs_count": (i32, [vp, vp, i64, vp, i32, i32, i32, vp, i64, vp]),
    "wf_sort_noalloc_max_pairs": (i64, []),
    "wf_sort_pairs_noalloc": (i32, [vp, vp, i64, i32, i32, i32, vp]),
    "wf_reorder": (i32, [vp, vp, vp, vp, vp, vp, i64, vp, pf]),
    "wf_scene_create": (i32, [vp, i64, i32, vp, i64, pi32, i32, vp, vp, C.POINTER(vp)]),
    "wf_scene_destroy": (i32, [vp]),
    "wf_scene_set_materials": (i32, [vp, vp, vp, i32, vp]),
    "wf_scene_get_materials": (i32, [vp, C.POINTER(i32), vp, vp]),
    "wf_scene_wide_info": (i32, [vp, pi64, C.POINTER(i32), pf]),
    "wf_scene_bounds": (i32, [vp, pf, pf]),
    "wf_scene_info": (i32, [vp, pi64, pi64, pi64, C.POINTER(i32)]),
    "wf_scene_download": (i32, [vp, pf, pi32]),
    "wf_trace": (i32, [vp, vp, vp, vp, i64, i32, vp, pf]),
    "wf_reorder_trace": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, i64, i32, vp, pf]),
    "wf_trace_stats": (i32, [vp, vp, vp, i64, i32, vp, vp, vp]),
    "wf_trace_sort": (i32, [vp, C.POINTER(WfRayBuffer), i32, i32, pf, pf, i32, vp, pf]),
    "wf_init_seeds": (i32, [vp, i64, i32, vp, pf]),
    "wf_raygen_primary": (i32, [pf, pf, vp, vp, i32, i32, i32, f32, vp, pf]),
    "wf_raygen_shadow": (i32, [vp, vp, vp, vp, vp, vp, vp, i64, i64, i32, pf, f32, vp, pf]),
    "wf_shade_extend": (i32, [vp, vp, vp, vp, vp, i64, vp, vp, vp, vp, i32, i32, i32, vp, i64,
                              pu32, vp, pf]),
    "wf_compact_scratch_bytes": (i64, [i64]),
    "wf_fill_f32": (i32, [vp, i64, f32, vp, pf]),
    "wf_reconstruct_smooth": (i32, [vp, vp, vp, vp, vp, vp, i64, pf, pf, i32, vp, pf]),
    "wf_reconstruct_shadow": (i32, [vp, vp, vp, vp, vp, i64, i64, i32, i32, vp, pf]),
    "wf_interpolate_colors": (i32, [vp, vp, i64, i32, f32, f32, vp, pf]),
    "wf_count_hits": (i32, [vp, i64, vp, vp, pf]),
    "wf_set_tuning": (i32, [cs, i32]),
    "wf_get_tuning": (i32, [cs, C.POINTER(i32)]),
    "wf_pixel_table": (i32, [i32, i32, pi32, pi32]),
    "wf_camera_matrices": (i32, [pf, pf, f32, f32, f32, f32, pf, pf, pf]),
    "wf_renderer_create": (i32, [cs, C.POINTER(vp)]),
    "wf_renderer_create_from_string": (i32, [cs, C.POINTER(vp)]),
    "wf_renderer_destroy": (i32, [vp]),
    "wf_renderer_build_scene": (i32, [vp]),
    "wf_renderer_set_asset_path": (i32, [vp, cs]),
    "wf_renderer_set_scene": (i32, [vp, vp, pf, pf]),
    "wf_renderer_get_scene": (i32, [vp, C.POINTER(vp)]),
    "wf_renderer_render": (i32, [vp, pf]),
    "wf_renderer_reset_frame_index": (i32, [vp]),
    "wf_renderer_read_pixels": (i32, [vp, i32, pf]),
    "wf_renderer_pixels_device": (i32, [vp, i32, C.POINTER(vp)]),
    "wf_renderer_save_image": (i32, [vp, cs, i32]),
    "wf_write_png": (i32, [cs, pf, i32, i32]),
    "wf_write_pfm": (i32, [cs, pf, i32, i32]),
    "wf_renderer_adjoint": (i32, [vp, i32, vp, vp, vp, vp, pf]),
    "wf_renderer_size": (i32, [vp, C.POINTER(i32), C.POINTER(i32)]),
    "wf_renderer_stats": (i32, [vp, pu64, pf, pf, pu64]),
    "wf_renderer_sort_logs": (i32, [vp, i32, vp, i32, C.POINTER(i32)]),
    "wf_renderer_set_decision_log": (i32, [vp, i32]),
    "wf_renderer_decisions": (i32, [vp, vp, vp, vp]),
    "wf_env_get_int": (i32, [vp, cs, C.POINTER(i32)]),
    "wf_env_get_float": (i32, [vp, cs, pf]),
    "wf_env_get_string": (i32, [vp, cs, C.c_char_p, i32]),
    "wf_env_get_vector": (i32, [vp, cs, pf]),
    "wf_renderer_set_capture": (i32, [vp, i32]),
    "wf_renderer_captured": (i32, [vp, i32, C.POINTER(vp), C.POINTER(vp), pi64]),
    "wf_procedural_boxes": (i32, [u64, i32, pi64, pi64, C.POINTER(i32), C.POINTER(i32), pf, pu32,
                                  pi32, pi32, pf, pf, pf, pf]),
    "wf_batch_create": (i32, [C.POINTER(vp), i32, vp, i32, i32, i32, C.POINTER(vp)]),
    "wf_batch_destroy": (i32, [vp]),
    "wf_batch_set_tuning": (i32, [vp, i32, i32, i32]),
    "wf_batch_get_tuning": (i32, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
    "wf_batch_autotune": (i32, [vp, pf]),
    "wf_batch_set_graph": (i32, [vp, i32]),
    "wf_batch_render": (i32, [vp, pf]),
    "wf_batch_read_pixels": (i32, [vp, i32, pf]),
    "wf_batch_stats": (i32, [vp, pu64, pi64]),
    "wf_procedural_scene": (i32, [i32, u64, i32, pi64, pi64, C.POINTER(i32), C.POINTER(i32), pf, pu32,
                                  pi32, pi32, pf, pf, pf, pf]),
    "wf_gltf_load": (i32, [cs, pi64, pi64, C.POINTER(i32), C.POINTER(i32), pf, pu32, pi32, pi32, pf,
                           pf, pf, pf]),
}

_lib = None


class WfError(RuntimeError):
    def __init__(self, fn: str, status: int, msg: str):
        super().__init__(f"{fn} failed with status {status}: {msg}")
        self.status = status


def load() -> C.CDLL:
    """Loads libwfpt.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue  # reported by missing_symbols()
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def missing_symbols() -> list[str]:
    """Symbols of include/wfpt.h that the built library does not export."""
    lib = load()
    return [n for n in SIGNATURES if getattr(lib, n, None) is None]


def check(fn_name: str, status: int) -> int:
    """Raises WfError for a negative status, with the library's message."""
    if status < 0:
        raise WfError(fn_name, status, load().wf_last_error().decode())
    return status


def call(name: str, *args):
    return check(name, getattr(load(), name)(*args))


_hip = None


def memcpy_d2d(dst: int, src: int, nbytes: int) -> None:
    """Synchronous device copy between raw device addresses (hipMemcpy)."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemcpy.restype = C.c_int
    rc = _hip.hipMemcpy(C.c_void_p(dst), C.c_void_p(src), C.c_size_t(nbytes), 3)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed ({rc})")


def f3(v) -> C.Array:
    return (C.c_float * 3)(*[float(x) for x in v])


def fptr(arr) -> C.POINTER(C.c_float):
    return C.cast(C.c_void_p(arr), pf) if isinstance(arr, int) else arr
