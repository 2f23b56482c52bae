"""The oracle's restatement of CSS-Python ``compact_fd6`` against the reference's own outputs.

``tests/golden/compact_fd6_ref.npz`` holds the reference function's results
(CSS-Python/derivatives.py:32-185 with linalg.py's Thomas solve), written by
``tests/golden/gen_compact_fd6.py`` for n in 5..128, ibc 0-5, dtype 0-3, with and without a
preloaded ``darr``.  These vectors pin the Python operator's halo-row ``dli/60`` scaling
(derivatives.py:91,141,177,181), the ``ibc == 5`` convention (``dli = n``) and the ``darr``
side effect at every length, where the Fortran rows fix them only at n = 61.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "compact_fd6_ref.npz")


_CASES = []


def golden_cases():
    if _CASES:
        return _CASES
    with np.load(GOLD) as f:
        z = {k: f[k] for k in f.files}
    off = z["off"]
    out = []
    for k in range(len(z["n"])):
        sl = slice(off[k], off[k + 1])
        out.append({"n": int(z["n"][k]), "ibc": int(z["ibc"][k]), "dtype": int(z["dtype"][k]),
                    "pre": int(z["pre"][k]), "dri": float(z["dri"][k]), "arr": z["arr"][sl],
                    "darr_in": z["darr_in"][sl], "result": z["result"][sl], "darr_out": z["darr_out"][sl],
                    "b1": z["b1"][4 * k:4 * k + 4], "b2": z["b2"][4 * k:4 * k + 4]})
    _CASES.extend(out)
    return _CASES


def compare(got, ref, tol):
    """Relative max error for finite references; identical +-inf pattern where the reference's
    Thomas solve divides by a zero pivot (n = 5 with Dirichlet-type closures at both walls)."""
    fin = np.isfinite(ref)
    if not fin.all():
        assert np.array_equal(np.isfinite(got), fin)
        assert np.array_equal(np.sign(got[~fin]), np.sign(ref[~fin]))
        return 0.0
    err = np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300)
    assert err < tol, err
    return err


def test_fixture_covers_the_grid():
    cs = golden_cases()
    assert len(cs) == 13 * 6 * 4 * 2
    assert {c["n"] for c in cs} >= {5, 8, 9, 17, 40, 60, 61, 64, 100, 128}
    # the fixture is not degenerate: the halo scaling matters away from n = 61
    c40 = [c for c in cs if c["n"] == 40 and c["dtype"] == 3 and c["ibc"] == 1 and c["pre"] == 0][0]
    ref_f = oracle.deriv_pencil(1, c40["arr"], 0, 0, 1, c40["b1"], c40["b2"], scale=39 * c40["dri"])
    assert np.abs(ref_f - c40["result"]).max() > 1e-3 * np.abs(ref_f).max()


@pytest.mark.parametrize("n", [5, 6, 7, 8, 9, 12, 17, 40, 60, 61, 64, 100, 128])
def test_oracle_matches_reference_vectors(n):
    for c in golden_cases():
        if c["n"] != n:
            continue
        darr = c["darr_in"].copy() if c["pre"] else None
        b1 = c["b1"] if c["dtype"] in (1, 3) else 0.0
        b2 = c["b2"] if c["dtype"] in (2, 3) else 0.0
        res, d_after = oracle.compact_fd6_py(c["dri"], c["arr"], b1, b2, c["ibc"], c["dtype"], darr)
        compare(res, c["result"], 1e-13)
        if c["pre"]:
            compare(d_after, c["darr_out"], 1e-13)
