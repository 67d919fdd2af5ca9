"""GPU: rmesh_sort_pairs_noalloc (device buffers, fixed per-device scratch,
no allocation) against rmesh_sort_pairs (host buffers, allocating) through
the C ABI: bit-identical keys and values on two calls in a row, the second
with a smaller input, at an even and an odd pass count; sizes past the
scratch capacity are refused."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pairs(rng, n, bits):
    keys = rng.integers(0, 1 << 63, n, dtype=np.uint64)
    # heavy ties on the sorted bits (stability is visible in the values)
    low = rng.integers(0, 1 << min(bits, 12), n, dtype=np.uint64)
    mask = np.uint64((1 << bits) - 1) if bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    keys = (keys & ~mask) | (low & mask)
    vals = np.arange(n, dtype=np.uint32)
    return keys, vals


@pytest.mark.parametrize("bits", [48, 40])  # 6 passes (result in place), 5 passes (copied back)
def test_noalloc_sort_matches_allocating_sort(bits):
    import torch

    from rmeshvk_amd import _lib
    from rmeshvk_amd.renderer import sort_pairs

    lib = _lib.load()
    rng = np.random.default_rng(bits)
    for n in (1_000_003, 70_001):  # second call smaller: nothing of the first may leak into it
        keys, vals = _pairs(rng, n, bits)
        want_k, want_v = sort_pairs(keys, vals, bits, device_id=0)
        dk = torch.from_numpy(keys.view(np.int64)).to("cuda:0")
        dv = torch.from_numpy(vals.view(np.int32)).to("cuda:0")
        torch.cuda.synchronize()
        st = lib.rmesh_sort_pairs_noalloc(ctypes.c_void_p(dk.data_ptr()), ctypes.c_void_p(dv.data_ptr()), n, bits, 0,
                                          ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        _lib.check(st)
        torch.cuda.synchronize()
        got_k = dk.cpu().numpy().view(np.uint64)
        got_v = dv.cpu().numpy().view(np.uint32)
        np.testing.assert_array_equal(got_k, want_k)
        np.testing.assert_array_equal(got_v, want_v)


def test_noalloc_sort_refuses_oversize():
    from rmeshvk_amd import _lib

    lib = _lib.load()
    st = lib.rmesh_sort_pairs_noalloc(ctypes.c_void_p(16), ctypes.c_void_p(16), (1 << 20) + 1, 48, 0, None)
    assert st == _lib.RMESH_ERR_INVALID
    assert b"scratch capacity" in lib.rmesh_last_error()
