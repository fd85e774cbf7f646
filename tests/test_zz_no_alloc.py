"""nfh_sort_perm_i32_noalloc (the allocation-free argsort) against
nf_sort_perm_i32 through the C ABI: two calls in a row, the second on a
smaller input, ascending and descending, keys whose high digits all keys
share (passes the plain entry point skips), and the refusals."""

import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod(hip):
    import torch
    return torch


def _bufs(torch, n):
    return [torch.empty(n, dtype=torch.int32, device="cuda") for _ in range(4)]


def _plain(hip, torch, keys, desc):
    perm, t0, t1, tp = _bufs(torch, keys.numel())
    rc = hip.nf_sort_perm_i32(keys.data_ptr(), keys.numel(), desc, perm.data_ptr(),
                              t0.data_ptr(), t1.data_ptr(), tp.data_ptr())
    torch.cuda.synchronize()
    assert rc == 0
    return perm


def _noalloc(hip, torch, keys, desc):
    perm, t0, t1, tp = _bufs(torch, keys.numel())
    torch.cuda.synchronize()
    rc = hip.nfh_sort_perm_i32_noalloc(keys.data_ptr(), keys.numel(), desc, perm.data_ptr(),
                                       t0.data_ptr(), t1.data_ptr(), tp.data_ptr())
    assert rc == 0
    return perm


@pytest.mark.parametrize("desc", [0, 1])
def test_noalloc_matches_plain_two_calls(hip, torch_mod, desc):
    torch = torch_mod
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    big = torch.randint(-(2 ** 31), 2 ** 31, ((1 << 22) + 5,), generator=g, device="cuda",
                        dtype=torch.int64).to(torch.int32)
    small = torch.randint(0, 256, (1000,), generator=g, device="cuda", dtype=torch.int32)
    print("noalloc sort: %d then %d keys, descending=%d" % (big.numel(), small.numel(), desc),
          flush=True)
    got_big = _noalloc(hip, torch, big, desc)
    got_small = _noalloc(hip, torch, small, desc)
    assert torch.equal(got_big, _plain(hip, torch, big, desc))
    assert torch.equal(got_small, _plain(hip, torch, small, desc))
    ref = torch.sort(small, stable=True, descending=bool(desc)).indices.to(torch.int32)
    assert torch.equal(got_small, ref)


def test_noalloc_allocates_no_device_memory(hip, torch_mod):
    torch = torch_mod
    keys = torch.randint(-1000, 1000, (1 << 20,), device="cuda", dtype=torch.int32)
    perm, t0, t1, tp = _bufs(torch, keys.numel())
    args = (keys.data_ptr(), keys.numel(), 0, perm.data_ptr(), t0.data_ptr(), t1.data_ptr(),
            tp.data_ptr())
    assert hip.nfh_sort_perm_i32_noalloc(*args) == 0
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(3):
        assert hip.nfh_sort_perm_i32_noalloc(*args) == 0
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    assert torch.equal(perm, torch.sort(keys, stable=True).indices.to(torch.int32))


def test_noalloc_refusals(hip, torch_mod):
    import numpy as np
    torch = torch_mod
    d = _bufs(torch, 8)
    p = [x.data_ptr() for x in d]
    # more than 2^24 keys: refused before any buffer is touched
    assert hip.nfh_sort_perm_i32_noalloc(p[0], (1 << 24) + 1, 0, p[1], p[2], p[3], p[0]) == -3
    # host buffers would need staging
    h = np.arange(8, dtype=np.int32)
    assert hip.nfh_sort_perm_i32_noalloc(h.ctypes.data, 8, 0, p[1], p[2], p[3], p[0]) == -3
    assert hip.nfh_sort_perm_i32_noalloc(None, 8, 0, p[1], p[2], p[3], p[0]) == -1
    assert hip.nfh_sort_perm_i32_noalloc(p[0], 0, 0, p[1], p[2], p[3], p[0]) == 0
