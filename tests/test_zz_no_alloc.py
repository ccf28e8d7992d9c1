"""bspai_hip_csr_from_entries_noalloc (the allocation-free variant of the
GPU canonicalization) against bspai_hip_csr_from_entries, through the C ABI:
identical CSR bits, on two calls in a row (the second smaller, so stale
scratch from the first must not leak into it), the same error messages, and
its limits rejected with a clear error."""
import ctypes as C

import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu

MAX_ENTRIES = 1 << 19


def _lib():
    import bspai_hip

    L = bspai_hip.load()
    fn = L.bspai_hip_csr_from_entries_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p,
                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)]
    return L, fn


def _entries(seed, rows, cols, count, dev):
    import torch

    rng = np.random.default_rng(seed)
    r = rng.integers(0, rows, count)
    c = rng.integers(0, cols, count)
    v = rng.standard_normal(count)
    # duplicates that cancel exactly and plain duplicates
    k = count // 10
    r[-k:], c[-k:] = r[:k], c[:k]
    v[-k // 2:] = -v[:k // 2]
    return tuple(torch.tensor(x, device=dev) for x in (r, c, v))


def _noalloc(ctx, fn, rows, cols, r, c, v):
    import torch

    dev = v.device
    n = v.numel()
    ptr = torch.full((rows + 1,), -7, dtype=torch.int32, device=dev)
    oc = torch.full((max(n, 1),), -7, dtype=torch.int32, device=dev)
    ov = torch.full((max(n, 1),), float("nan"), dtype=torch.float64, device=dev)
    nnz = C.c_int64(-1)
    torch.cuda.synchronize()
    rc = fn(ctx.h, rows, cols, n, r.data_ptr(), c.data_ptr(), v.data_ptr(), ptr.data_ptr(),
            oc.data_ptr(), ov.data_ptr(), C.byref(nnz))
    return rc, ptr, oc[:nnz.value], ov[:nnz.value]


def _same(a, b):
    assert np.array_equal(a[0].cpu().numpy(), b[0].cpu().numpy())
    assert np.array_equal(a[1].cpu().numpy(), b[1].cpu().numpy())
    assert np.array_equal(bits(a[2].cpu().numpy()), bits(b[2].cpu().numpy()))


def test_noalloc_matches_allocating_entry_point(ctx):
    import torch

    import bspai_hip

    L, fn = _lib()
    dev = torch.device("cuda:0")
    # first call: near the limit; second: a different, smaller matrix
    for seed, rows, cols, count in ((1, 40000, 30000, MAX_ENTRIES), (2, 700, 900, 5000)):
        r, c, v = _entries(seed, rows, cols, count, dev)
        want = bspai_hip.csr_from_entries(rows, cols, r, c, v, ctx)
        rc, ptr, oc, ov = _noalloc(ctx, fn, rows, cols, r, c, v)
        assert rc == 0, L.bspai_hip_last_error(ctx.h)
        _same((ptr, oc, ov), want)
    # an empty input gives empty rows
    rc, ptr, oc, ov = _noalloc(ctx, fn, 5, 5, *(torch.empty(0, dtype=t, device=dev)
                                                 for t in (torch.int64, torch.int64,
                                                           torch.float64)))
    assert rc == 0 and ptr.cpu().tolist() == [0] * 6 and oc.numel() == 0


def test_noalloc_errors_and_limits(ctx):
    import torch

    L, fn = _lib()
    dev = torch.device("cuda:0")
    t = lambda x, dt: torch.tensor(x, dtype=dt, device=dev)  # noqa: E731
    cases = [
        (([0, 2], [0, 0], [1.0, 2.0]), "Entry index outside matrix dimensions"),
        (([0, 1], [0, 0], [1.0, float("inf")]), "Matrix entries must be finite"),
        (([0, 0], [1, 1], [1e308, 1e308]), "Duplicate sum is nonfinite"),
    ]
    for (r, c, v), msg in cases:
        rc, *_ = _noalloc(ctx, fn, 2, 2, t(r, torch.int64), t(c, torch.int64),
                          t(v, torch.float64))
        assert rc == 1 and L.bspai_hip_last_error(ctx.h).decode() == msg
    # beyond the limit: a clear error, nothing launched
    big = torch.zeros(MAX_ENTRIES + 1, dtype=torch.int64, device=dev)
    rc, *_ = _noalloc(ctx, fn, 10, 10, big, big, torch.ones(MAX_ENTRIES + 1,
                                                            dtype=torch.float64, device=dev))
    assert rc == 1 and "at most" in L.bspai_hip_last_error(ctx.h).decode()
    # host arrays are refused (no staging copies)
    hr = np.zeros(3, np.int64)
    hv = np.ones(3)
    ptr = torch.empty(3, dtype=torch.int32, device=dev)
    oc = torch.empty(3, dtype=torch.int32, device=dev)
    ov = torch.empty(3, dtype=torch.float64, device=dev)
    nnz = C.c_int64()
    rc = fn(ctx.h, 2, 2, 3, hr.ctypes.data, hr.ctypes.data, hv.ctypes.data, ptr.data_ptr(),
            oc.data_ptr(), ov.data_ptr(), C.byref(nnz))
    assert rc == 1 and "device arrays" in L.bspai_hip_last_error(ctx.h).decode()
