"""The GPU drop-in ``cssx.derivatives.compact_fd6`` against the reference's own outputs.

The vectors in ``tests/golden/compact_fd6_ref.npz`` are CSS-Python's ``compact_fd6``
(derivatives.py:32-185) evaluated by ``tests/golden/gen_compact_fd6.py``: n = 5..128,
ibc 0-5, dtype 0-3, with and without a preloaded ``darr``.  Tolerance: 1e-12 of the
result's max norm (the kernels' segmented solve reorders the Thomas arithmetic; rounding
only); the one singular system (n = 5, ibc = 1, dtype = 0) must give the same +-inf pattern.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from test_golden_compact_fd6 import compare, golden_cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [5, 6, 7, 8, 9, 12, 17, 40, 60, 61, 64, 100, 128])
def test_drop_in_matches_reference_vectors(n):
    import cssx.derivatives as D
    assert torch.cuda.is_available()
    seen = 0
    for c in golden_cases():
        if c["n"] != n:
            continue
        b1 = c["b1"] if c["dtype"] in (1, 3) else 0.0
        b2 = c["b2"] if c["dtype"] in (2, 3) else 0.0
        if c["pre"]:
            darr = c["darr_in"].copy()
            got = D.compact_fd6(c["dri"], c["arr"], b1, b2, c["ibc"], c["dtype"], darr)
            compare(darr, c["darr_out"], 1e-12)
        else:
            got = D.compact_fd6(c["dri"], c["arr"], b1, b2, c["ibc"], c["dtype"])
        compare(got, c["result"], 1e-12)
        seen += 1
    assert seen == 6 * 4 * 2


@pytest.mark.parametrize("n", [5, 7, 40])
def test_drop_in_batched_pencils_match_reference_vectors(n):
    """All cases of one (n, ibc, dtype) family stacked as pencils of one launch (axis 0)."""
    import cssx.derivatives as D
    for ibc in range(6):
        for dtype in range(4):
            cs = [c for c in golden_cases() if c["n"] == n and c["ibc"] == ibc and c["dtype"] == dtype]
            arr = np.stack([c["arr"] for c in cs], axis=1)              # (n, 2)
            darr = np.stack([c["darr_in"] for c in cs], axis=1)
            b1 = np.stack([c["b1"] for c in cs], axis=1) if dtype in (1, 3) else 0.0
            b2 = np.stack([c["b2"] for c in cs], axis=1) if dtype in (2, 3) else 0.0
            got = D.compact_fd6(cs[0]["dri"], arr, b1, b2, ibc, dtype, darr, axis=0)
            for j, c in enumerate(cs):
                compare(got[:, j], c["result"], 1e-12)
                if c["pre"]:
                    compare(darr[:, j], c["darr_out"], 1e-12)
