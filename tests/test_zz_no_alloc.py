"""ever_morton_order_noalloc (the Morton order without a workspace argument, on the library's own scratch array)
against ever_morton_order with a caller workspace: equal orders on two calls in a row on one GPU, the second with a
smaller input (the scratch keeps the first call's contents), including ties and a size at the limit."""
import ctypes

import numpy as np
import pytest
import torch

from ever_amd import _capi as C

pytestmark = pytest.mark.gpu
D = torch.device("cuda:0")


def _with_workspace(means):
    L = C.lib()
    n = means.shape[0]
    order = torch.empty(n, dtype=torch.int32, device=D)
    wsb = L.ever_morton_workspace_bytes(n)
    ws = torch.empty(wsb, dtype=torch.uint8, device=D)
    C.check(L.ever_morton_order(C.ptr(means), n, C.ptr(order), C.ptr(ws), wsb, C.stream_ptr()), "ever_morton_order")
    return order


def _no_alloc(means):
    L = C.lib()
    n = means.shape[0]
    order = torch.empty(n, dtype=torch.int32, device=D)
    C.check(L.ever_morton_order_noalloc(C.ptr(means), n, C.ptr(order), C.stream_ptr()), "ever_morton_order_noalloc")
    return order


def _means(n, seed):
    g = torch.Generator().manual_seed(seed)
    m = torch.randn((n, 3), generator=g) * 2.0
    m[: n // 10] = m[n // 10: 2 * (n // 10)]  # duplicated positions: ties keep index order
    return m.to(D).contiguous()


def test_no_alloc_morton_order_equals_workspace_form_on_two_calls():
    print("case: noalloc Morton order, 2^20 then 777 primitives", flush=True)
    for n, seed in ((C.EVER_MORTON_NOALLOC_MAX_PRIMS, 1), (777, 2)):
        m = _means(n, seed)
        a = _no_alloc(m)
        b = _with_workspace(m)
        torch.cuda.synchronize()
        assert torch.equal(a, b), n
        assert torch.equal(torch.sort(a.long()).values, torch.arange(n, device=D))  # a permutation
    # the size limit is enforced before any GPU work
    L = C.lib()
    big = C.EVER_MORTON_NOALLOC_MAX_PRIMS + 1
    assert L.ever_morton_order_noalloc(ctypes.c_void_p(1), big, ctypes.c_void_p(1), C.stream_ptr()) == C.EVER_ERR_INVALID
