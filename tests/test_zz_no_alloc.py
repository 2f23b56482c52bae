"""cssx_compact_fd6_noalloc, the variant of cssx_compact_fd6 that allocates nothing: its plan
table lives in the library's fixed per-device scratch array.  Every call is compared bit for
bit with cssx_compact_fd6 on the same inputs (the same kernels read the same FP64 plan), over
calls in a row whose sizes shrink, so that each call overwrites what the previous one left in
the shared scratch."""
import ctypes

import numpy as np
import pytest
import torch

from cssx import _lib

pytestmark = pytest.mark.gpu


def _fn(L, name):
    fn = getattr(L, name)
    fn.argtypes = _lib.SIGNATURES["cssx_compact_fd6"][0]
    fn.restype = ctypes.c_int
    return fn


def _run(fn, n, npen, dri, ibc, dtype, arr, b1, b2, darr):
    res = torch.full((n, npen), float("nan"), dtype=torch.float64, device="cuda")
    d = darr.clone()
    stream = ctypes.c_void_p(int(torch.cuda.current_stream().cuda_stream))
    rc = fn(n, npen, float(dri), ibc, dtype, arr.data_ptr(), b1.data_ptr(), b2.data_ptr(), d.data_ptr(),
            res.data_ptr(), stream)
    torch.cuda.synchronize()
    return rc, res, d


# (n, pencils): each call smaller than the one before; n = 6 takes the short-pencil kernel
SIZES = [(1024, 3000), (700, 1500), (257, 900), (64, 400), (9, 100), (6, 50)]


def test_noalloc_matches_compact_fd6():
    assert torch.cuda.is_available()
    L = _lib.lib()
    ref_fn, na_fn = _fn(L, "cssx_compact_fd6"), _fn(L, "cssx_compact_fd6_noalloc")
    rng = np.random.default_rng(77)
    for n, npen in SIZES:
        for ibc, dtype in ((1, 0), (4, 0), (5, 2), (0, 3), (2, 1)):
            print(f"noalloc n={n} pencils={npen} ibc={ibc} dtype={dtype}", flush=True)
            arr = torch.as_tensor(rng.standard_normal((n, npen)), device="cuda")
            b1 = torch.as_tensor(rng.standard_normal((4, npen)), device="cuda")
            b2 = torch.as_tensor(rng.standard_normal((4, npen)), device="cuda")
            darr = torch.zeros((n, npen), dtype=torch.float64, device="cuda")
            darr[0] = torch.as_tensor(rng.standard_normal(npen), device="cuda")
            darr[-1] = torch.as_tensor(rng.standard_normal(npen), device="cuda")
            dri = 1.0 / 0.25
            rc0, r0, d0 = _run(ref_fn, n, npen, dri, ibc, dtype, arr, b1, b2, darr)
            rc1, r1, d1 = _run(na_fn, n, npen, dri, ibc, dtype, arr, b1, b2, darr)
            assert rc0 == 0 and rc1 == 0, (rc0, rc1)
            assert torch.isfinite(r0).all()
            assert torch.equal(r1, r0), (n, npen, ibc, dtype, (r1 - r0).abs().max().item())
            assert torch.equal(d1, d0), (n, npen, ibc, dtype)


def test_noalloc_rejects_pencils_beyond_its_scratch():
    """n > 1024 has no plan on either entry point (CSSX_EINVAL, nothing launched)."""
    L = _lib.lib()
    na_fn = _fn(L, "cssx_compact_fd6_noalloc")
    n, npen = 1025, 4
    arr = torch.zeros((n, npen), dtype=torch.float64, device="cuda")
    res = torch.zeros_like(arr)
    rc = na_fn(n, npen, 1.0, 1, 0, arr.data_ptr(), None, None, None, res.data_ptr(), None)
    assert rc == _lib.CSSX_EINVAL
