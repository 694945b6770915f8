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
