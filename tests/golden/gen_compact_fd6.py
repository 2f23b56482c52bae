#!/usr/bin/env python3
"""Write golden vectors of CSS-Python's ``compact_fd6`` (derivatives.py:32-185).

Runs only where the reference tree is present (never on a machine without it): it imports
``CSS-Python/derivatives.py`` and ``linalg.py`` from the reference checkout given by
``--ref`` (default ``/root/reference``), evaluates them on seeded inputs and stores inputs and
outputs in ``compact_fd6_ref.npz`` beside this script.  Nothing of the reference is written
but its outputs.

Cases: n in N_LIST, ibc 0-5, dtype 0-3, each with ``darr=None`` and with a preloaded ``darr``
(Neumann values f'/dri in rows 0 and n-1, noise elsewhere, which every assembled row
overwrites).  The reference tests ``darr == None``, which raises for a plain ndarray, so the
preloaded case passes an ndarray subclass whose comparison with None is False; every other
operation on it is the ndarray's own.

Stored per case k (arrays concatenated, ``off[k]:off[k+1]`` slices the n-long ones):
``arr``, ``darr_in`` (zeros for the None case), ``result``, ``darr_out`` (what the reference
leaves in the caller's darr; zeros for the None case), ``b1``/``b2`` (4 each), and the
scalars ``n``, ``ibc``, ``dtype``, ``pre`` (1 = preloaded darr), ``dri``.

  python tests/golden/gen_compact_fd6.py [--ref /root/reference]
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

N_LIST = (5, 6, 7, 8, 9, 12, 17, 40, 60, 61, 64, 100, 128)
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "compact_fd6_ref.npz")


class _Preloaded(np.ndarray):
    """ndarray whose ``== None`` is False (the reference's ``if (darr == None)``)."""

    def __eq__(self, other):
        if other is None:
            return False
        return np.ndarray.__eq__(self, other)

    def __ne__(self, other):
        if other is None:
            return True
        return np.ndarray.__ne__(self, other)


def _inputs(n, ibc, dtype, pre):
    rng = np.random.default_rng(1000 * n + 100 * ibc + 10 * dtype + pre)
    x1, x2 = 0.72, 0.97
    dri = 1.0 / (x2 - x1)
    arr = rng.standard_normal(n)
    b1 = rng.standard_normal(4)
    b2 = rng.standard_normal(4)
    darr = np.zeros(n)
    if pre:
        darr = rng.standard_normal(n)
        darr[0] = rng.standard_normal() / dri
        darr[-1] = rng.standard_normal() / dri
    return dri, arr, b1, b2, darr


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    a = ap.parse_args()
    sys.path.insert(0, os.path.join(a.ref, "CSS-Python"))
    import derivatives as ref  # the reference module, read from the reference tree

    cols = {k: [] for k in ("arr", "darr_in", "result", "darr_out", "b1", "b2")}
    meta = {k: [] for k in ("n", "ibc", "dtype", "pre", "dri")}
    for n in N_LIST:
        for ibc in range(6):
            for dtype in range(4):
                for pre in (0, 1):
                    dri, arr, b1, b2, darr = _inputs(n, ibc, dtype, pre)
                    d_in = darr.copy()
                    if pre:
                        dd = darr.copy().view(_Preloaded)
                        res = ref.compact_fd6(dri, arr.copy(), b1.copy(), b2.copy(), ibc, dtype, dd)
                        d_out = np.asarray(dd).copy()
                    else:
                        res = ref.compact_fd6(dri, arr.copy(), b1.copy(), b2.copy(), ibc, dtype)
                        d_out = np.zeros(n)
                    for k, v in (("arr", arr), ("darr_in", d_in), ("result", np.asarray(res, dtype=np.float64)),
                                 ("darr_out", d_out), ("b1", b1), ("b2", b2)):
                        cols[k].append(v)
                    for k, v in (("n", n), ("ibc", ibc), ("dtype", dtype), ("pre", pre), ("dri", dri)):
                        meta[k].append(v)
    ns = np.array(meta["n"], dtype=np.int64)
    off = np.concatenate([[0], np.cumsum(ns)])
    np.savez_compressed(OUT, off=off, **{k: np.concatenate(v) for k, v in cols.items()},
                        **{k: np.array(v) for k, v in meta.items()})
    print(f"wrote {OUT}: {len(ns)} cases, n in {N_LIST}")


if __name__ == "__main__":
    main()
