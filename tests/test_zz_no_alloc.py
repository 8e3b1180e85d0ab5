"""drx_crt_noalloc (the allocation-free CRT) against drx_crt through the C
ABI: two calls in a row, the second on fewer and narrower rows than the
first (the shared staging keeps the last call's contents), on both Garner
kernels, plus the size limit."""

import numpy as np
import pytest

import dreams_rns as D
from dreams_rns import _native as N

pytestmark = pytest.mark.gpu


def _call(ctx, fn, res, K, stride):
    n = res.shape[0]
    r = N.pinned_zeros((n, stride), np.int64)
    r[:, :K] = res
    limbs = N.pinned_zeros((n, stride), np.uint32)
    neg = N.pinned_zeros((n,), np.uint8)
    limbs[...] = 0xDEADBEEF  # every limb must be written
    neg[...] = 7
    ctx._check(getattr(ctx.lib, fn)(ctx.h, r.ctypes.data, n, K, stride, limbs.ctypes.data, neg.ctypes.data))
    return limbs.copy(), neg.copy()


def _rows(ctx, n, K, seed):
    ctx.ensure_primes(K)
    rng = np.random.default_rng(seed)
    pr = ctx.primes[:K]
    res = (rng.integers(0, 1 << 62, size=(n, K), dtype=np.int64) % pr).astype(np.int64)
    res[0] = 0
    res[1] = 1
    return res


@pytest.mark.parametrize("first,second", [((300, 64, 70), (17, 32, 33)), ((40, 2001, 2001), (5, 300, 301))])
def test_noalloc_equals_drx_crt_twice(gpu_ctx, first, second):
    for i, (n, K, stride) in enumerate((first, second)):
        res = _rows(gpu_ctx, n, K, 100 * K + i)
        print("noalloc crt call", i, "n", n, "K", K, "stride", stride, flush=True)
        want = _call(gpu_ctx, "drx_crt", res, K, stride)
        got = _call(gpu_ctx, "drx_crt_noalloc", res, K, stride)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), (i, n, K)
        assert not got[0][:, K:].any()


def test_noalloc_refuses_more_than_its_scratch(gpu_ctx):
    gpu_ctx.ensure_primes(32)
    n, K = 30000, 32  # 960000 residues: above the 8 MiB staging
    res = np.zeros((n, K), dtype=np.int64)
    limbs = np.zeros((n, K), dtype=np.uint32)
    neg = np.zeros(n, dtype=np.uint8)
    rc = gpu_ctx.lib.drx_crt_noalloc(gpu_ctx.h, res.ctypes.data, n, K, K, limbs.ctypes.data, neg.ctypes.data)
    assert rc == -4 and b"staging" in gpu_ctx.lib.drx_last_error(gpu_ctx.h)
    # and the CRT of a single value through the public helper still works afterwards
    assert D.crt_reconstruct([2, 3], [5, 7]) == (17, 35)
