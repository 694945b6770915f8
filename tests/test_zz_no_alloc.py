"""The allocation-free sort (wf_sort_pairs_noalloc, scratch in a device
array inside the library) against wf_sort_pairs with a caller scratch, bit
for bit, on two and three calls in a row with shrinking inputs (the array
keeps the previous call's contents), for u64 and u32 keys, and its size
limit."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _keys(n, key_bytes, bits, seed):
    rng = np.random.default_rng(seed)
    dt = np.uint64 if key_bytes == 8 else np.uint32
    k = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(dt)
    k[: n // 4] = k[n // 2: n // 2 + n // 4]          # duplicates: stability matters
    return k


@pytest.mark.parametrize("key_bytes,bits,digit_bits", [(8, 60, 8), (4, 30, 11), (8, 64, 4)])
def test_noalloc_sort_matches_sort_pairs(dev, key_bytes, bits, digit_bits):
    from vulkanwavefrontpathtracer_amd import _capi

    lib = _capi.load()
    assert lib.wf_sort_noalloc_max_pairs() == 1 << 20
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for call_no, n in enumerate((1 << 20, 70_001, 3_000)):
        print(f"[noalloc] call {call_no}: n={n} key_bytes={key_bytes} bits={bits} "
              f"digits={digit_bits}", flush=True)
        k = _keys(n, key_bytes, bits, 100 * call_no + bits)
        td = torch.int64 if key_bytes == 8 else torch.int32
        ka = torch.from_numpy(k.view(np.int64 if key_bytes == 8 else np.int32)).to(dev)
        va = torch.arange(n, dtype=torch.int32, device=dev)
        kb, vb = ka.clone(), va.clone()
        need = lib.wf_sort_scratch_bytes(C.c_int64(n), key_bytes, bits, digit_bits)
        scr = torch.empty(need, dtype=torch.uint8, device=dev)
        _capi.call("wf_sort_pairs", C.c_void_p(ka.data_ptr()), C.c_void_p(va.data_ptr()),
                   C.c_int64(n), key_bytes, bits, digit_bits, C.c_void_p(scr.data_ptr()),
                   C.c_int64(need), stream, None)
        _capi.call("wf_sort_pairs_noalloc", C.c_void_p(kb.data_ptr()), C.c_void_p(vb.data_ptr()),
                   C.c_int64(n), key_bytes, bits, digit_bits, stream)
        torch.cuda.synchronize()
        assert torch.equal(ka, kb) and torch.equal(va, vb), (call_no, n)
        order = np.argsort(k, kind="stable")
        assert np.array_equal(vb.cpu().numpy(), order.astype(np.int32))
        assert td == ka.dtype


def test_noalloc_sort_rejects_beyond_limit(dev):
    from vulkanwavefrontpathtracer_amd import _capi

    lib = _capi.load()
    n = (1 << 20) + 1
    k = torch.zeros(n, dtype=torch.int64, device=dev)     # buffers that would fit, unused
    v = torch.zeros(n, dtype=torch.int32, device=dev)
    st = lib.wf_sort_pairs_noalloc(C.c_void_p(k.data_ptr()), C.c_void_p(v.data_ptr()),
                                   C.c_int64(n), 8, 60, 8, None)
    assert st == -5
    assert b"allocation-free limit" in lib.wf_last_error()
    assert lib.wf_sort_pairs_noalloc(None, None, C.c_int64(1), 8, 60, 3, None) == -1
