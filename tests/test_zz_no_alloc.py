"""The allocation-free sort entry point (sphx_sort_keys_noalloc) against sphx_sort_keys.

Called through the C ABI on cuda:0 twice in a row, the second call smaller than the first, so
the fixed scratch holds the first call's contents when the second runs. Keys are drawn with many
repeats, so the order also checks that equal keys keep their original order (a stable sort).
"""
import ctypes as C

import numpy as np
import pytest

from sphexa_amd import abi

pytestmark = pytest.mark.gpu

MAX_KEYS = 1 << 20  # SPHX_SORT_NOALLOC_MAX_KEYS


def _keys(n, seed):
    import torch

    rng = np.random.default_rng(seed)
    # 63-bit Hilbert-range keys with repeats: a few thousand distinct values, plus removeKey (2^63)
    base = rng.integers(0, 1 << 63, size=4096, dtype=np.uint64)
    k = base[rng.integers(0, base.size, size=n)]
    k[rng.integers(0, n, size=max(1, n // 100))] = np.uint64(1) << np.uint64(63)
    return torch.from_numpy(k.view(np.int64).copy()).cuda()


def _noalloc(lib, keys, order, stream):
    return lib.sphx_sort_keys_noalloc(0, C.c_void_p(stream), C.c_void_p(keys.data_ptr()),
                                      C.c_void_p(order.data_ptr()), keys.numel())


def test_noalloc_sort_matches_workspace_sort():
    import torch

    from sphexa_amd.lib import Workspace, load

    lib = load()
    ws = Workspace(0)
    stream = torch.cuda.current_stream().cuda_stream
    for n, seed in ((MAX_KEYS, 3), (200_003, 4)):
        print(f"noalloc sort: n={n} seed={seed}", flush=True)
        keys = _keys(n, seed)
        ref_keys, ref_order = keys.clone(), torch.empty(n, dtype=torch.int32, device="cuda")
        ws.sort_keys(ref_keys, ref_order)
        got_keys, got_order = keys.clone(), torch.full((n,), -1, dtype=torch.int32, device="cuda")
        code = _noalloc(lib, got_keys, got_order, stream)
        assert code == abi.SPHX_OK, lib.sphx_last_error(None).decode()
        torch.cuda.synchronize()
        assert torch.equal(got_keys, ref_keys)
        assert torch.equal(got_order, ref_order)
        # and the order is the stable one
        host = keys.cpu().numpy().view(np.uint64)
        assert np.array_equal(got_order.cpu().numpy(), np.argsort(host, kind="stable").astype(np.int32))
    ws.close()


def test_noalloc_sort_rejects_more_than_the_maximum():
    import torch

    from sphexa_amd.lib import load

    lib = load()
    n = MAX_KEYS + 1
    keys = torch.zeros(n, dtype=torch.int64, device="cuda")
    order = torch.zeros(n, dtype=torch.int32, device="cuda")
    code = _noalloc(lib, keys, order, torch.cuda.current_stream().cuda_stream)
    assert code == abi.SPHX_ERR_ARG
    msg = lib.sphx_last_error(None).decode()
    assert "exceed the maximum" in msg and str(MAX_KEYS) in msg, msg
    # an empty sort is a no-op that succeeds
    assert _noalloc(lib, keys[:0], order[:0], torch.cuda.current_stream().cuda_stream) == abi.SPHX_OK
