"""cms_nearest_noalloc (the allocation-free nearest-point search) against
cms_nearest through the C ABI: identical squared distances and first
indices, bit for bit, in both precisions, on consecutive calls that share
the device scratch (largest allowed size first, then smaller inputs), with
ties from duplicated references; sizes beyond the scratch are refused."""

import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def points(nq, nr, seed):
    rng = np.random.default_rng(seed)
    q = rng.uniform(-30, 30, size=(nq, 3))
    r = rng.uniform(-30, 30, size=(nr, 3))
    k = min(nr // 4, 500)
    r[nr - k:] = r[:k]                  # exact duplicates: first-index ties
    m = min(nq, k) // 2
    q[:m] = r[:m]                       # zero distances
    return q, r


def search(ctx, name, prec, dq, dr, with_idx=True):
    import torch
    nq, nr = dq.shape[0], dr.shape[0]
    d2 = torch.full((max(nq, 1),), -1.0, dtype=torch.float64, device="cuda")
    idx = torch.full((max(nq, 1),), -7, dtype=torch.int32, device="cuda") if with_idx else None
    ctx(name, prec, ctypes.c_void_p(dq.data_ptr()), nq, ctypes.c_void_p(dr.data_ptr()), nr,
        ctypes.c_void_p(d2.data_ptr()), ctypes.c_void_p(idx.data_ptr()) if with_idx else None)
    return d2[:nq].cpu().numpy(), (idx[:nq].cpu().numpy() if with_idx else None)


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_matches_cms_nearest(prec):
    import torch
    from cms_cuda import _lib
    ctx = _lib.Context(0, torch.cuda.current_stream().cuda_stream)
    big = _lib.NOALLOC_MAX_POINTS
    for call, (nq, nr, seed) in enumerate([(big, big, 3), (20000, 15000, 4), (3000, 2500, 5), (700, 90000, 6)]):
        say("precision %d call %d: %d queries, %d references" % (prec, call, nq, nr))
        q, r = points(nq, nr, seed)
        dq = torch.from_numpy(q).cuda()
        dr = torch.from_numpy(r).cuda()
        ref_d2, ref_idx = search(ctx, "cms_nearest", prec, dq, dr)
        got_d2, got_idx = search(ctx, "cms_nearest_noalloc", prec, dq, dr)
        assert np.array_equal(got_d2, ref_d2), (prec, call)
        assert np.array_equal(got_idx, ref_idx), (prec, call)
        only_d2, _ = search(ctx, "cms_nearest_noalloc", prec, dq, dr, with_idx=False)
        assert np.array_equal(only_d2, ref_d2), (prec, call)
        if nq <= 3000:
            # brute force on the host for the small calls: first-index argmin
            from scipy.spatial.distance import cdist
            D = cdist(r, q, metric="sqeuclidean")
            bi = np.argmin(D, axis=0)
            assert np.array_equal(got_idx, bi)
            assert np.array_equal(got_d2, D[bi, np.arange(nq)])
    ctx.close()


def test_noalloc_limits():
    import torch
    from cms_cuda import _lib
    ctx = _lib.Context(0, torch.cuda.current_stream().cuda_stream)
    big = _lib.NOALLOC_MAX_POINTS
    dq = torch.zeros((big + 1, 3), dtype=torch.float64, device="cuda")
    dr = torch.ones((10, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="scratch limit"):
        search(ctx, "cms_nearest_noalloc", 0, dq, dr)
    with pytest.raises(ValueError, match="scratch limit"):
        search(ctx, "cms_nearest_noalloc", 1, dr, dq)
    # no queries: nothing to do; no references: the same as cms_nearest
    d2, idx = search(ctx, "cms_nearest_noalloc", 0, dq[:0], dr)
    assert d2.size == 0
    q = torch.rand((50, 3), dtype=torch.float64, device="cuda")
    a = search(ctx, "cms_nearest", 0, q, dr[:0])
    b = search(ctx, "cms_nearest_noalloc", 0, q, dr[:0])
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    ctx.close()
