"""sbp_apply_dot_noalloc (the entry point that allocates nothing: block
partials and results in the library's static device array) against
sbp_apply_dot through the C ABI: bit-identical products and inner products
on two calls in a row, the second on a smaller problem, a batch handle, and
the scratch-size check."""
import ctypes as C

import numpy as np
import pytest
import torch

from sc_2022_amd import _capi as capi

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def handle(N, p, nprob=1, **prm):
    lib = capi.load()
    h = C.c_void_p()
    params = capi.default_params(**prm)
    assert lib.sbp_create_batch(N, p, nprob, C.byref(params), C.byref(h)) == capi.SBP_OK
    return lib, h


def both(lib, h, N, nprob, seed):
    g = torch.Generator(DEV).manual_seed(seed)
    u = torch.randn(nprob * N * N, dtype=torch.float64, device=DEV, generator=g)
    Au1 = torch.empty_like(u)
    Au2 = torch.full_like(u, float("nan"))
    dot1 = (C.c_double * nprob)()
    dot2 = torch.full((nprob,), float("nan"), dtype=torch.float64, device=DEV)
    torch.cuda.synchronize()
    p = lambda t: C.c_void_p(t.data_ptr())
    assert lib.sbp_apply_dot(h, p(u), p(Au1), dot1) == capi.SBP_OK
    assert lib.sbp_apply_dot_noalloc(h, p(u), p(Au2), p(dot2)) == capi.SBP_OK
    torch.cuda.synchronize()
    assert torch.equal(Au1, Au2)
    d2 = dot2.cpu().numpy()
    assert np.array_equal(np.array(dot1[:]), d2), (dot1[:], d2)
    # and the inner products are u'Au
    ref = (u.view(nprob, -1) * Au1.view(nprob, -1)).sum(1).cpu().numpy()
    assert np.allclose(d2, ref, rtol=1e-12, atol=0)


def test_noalloc_matches_apply_dot_twice_then_smaller():
    lib, h1 = handle(1025, 4)
    lib, h2 = handle(257, 2)
    try:
        both(lib, h1, 1025, 1, 1)
        both(lib, h1, 1025, 1, 2)     # second call on the same GPU array
        both(lib, h2, 257, 1, 3)      # then a smaller problem
    finally:
        lib.sbp_destroy(h1)
        lib.sbp_destroy(h2)


def test_noalloc_batch_handle():
    lib, h = handle(129, 4, nprob=5)
    try:
        both(lib, h, 129, 5, 4)
    finally:
        lib.sbp_destroy(h)


def test_noalloc_rejects_more_partials_than_scratch():
    """64 x 4 tiles at 1025^2 give 17 * 257 = 4369 partials per problem; 31
    problems need 135 439 > 2^17 slots."""
    lib, h = handle(1025, 2, nprob=31, tile_i=64, tile_j=4)
    try:
        n = 31 * 1025 * 1025
        u = torch.zeros(n, dtype=torch.float64, device=DEV)
        Au = torch.empty_like(u)
        d = torch.empty(31, dtype=torch.float64, device=DEV)
        p = lambda t: C.c_void_p(t.data_ptr())
        assert lib.sbp_apply_dot_noalloc(h, p(u), p(Au), p(d)) == capi.SBP_ERR_SCRATCH
        assert b"scratch" in lib.sbp_status_string(capi.SBP_ERR_SCRATCH)
    finally:
        lib.sbp_destroy(h)
