"""Pins the wall values of the Laplacian cross-term sweeps (Laplacian_Cross_Terms_Radial/_Theta,
rhs_mpi.f90:2070-2296) in the oracle.

The sweeps (``oracle.lap_cross``, taken from inside the oracle's RHS) are, in continuous form,
with X1 = dw/dth / r, X2 = du/dph / (r sin), X3 = du/dth / r, X4 = dv/dph / (r sin),
X5 = dw/dph / (r sin):

    d(r X1)/dr = d2w/drdth        d(r X2)/dr = d2u/drdph / sin    d(r X3)/dr = d2u/drdth
    d(r X4)/dr = d2v/drdph / sin  d(sin X4)/dth = d2v/dthdph / r  d(sin X5)/dth = d2w/dthdph / r

Only the Neumann codes read wall values (w and v: stress-free r walls; v: the theta walls):
  (a) r sweep of X1 at r1: r1 X1/(r1 dxdr) (:2113-2117), the value a stress-free wall
      (d(w/r)/dr = 0) implies: d2w/drdth = d(w/r)/dth = X1;
  (b) r sweep of X1 at r2: never written, so 0;
  (c) r sweep of X4 at r1: never written, so 0;
  (d) r sweep of X4 at r2: r2 X4/(r2 dxdr), set under Bottom_Stress_Free (:2120-2125), again
      the stress-free value;
  (e) theta sweep of X4: sin X4 cot (:2243-2250), the value a stress-free theta wall
      (d(v/sin)/dth = 0) implies.
State A has w/r = (1 + cos pi xi) F (X1 != 0 at r1, 0 at r2) and v/r = (1 - cos pi xi) sin G
(X4 = 0 at r1, != 0 at r2), both stress free at every wall: all five rules are exact for it,
and every sweep converges to the continuous value up to and including the wall rows.  Any
other choice for (a), (d) or (e) leaves an O(1) error next to that wall.  State B swaps the
radial profiles: X1 != 0 at r2 and X4 != 0 at r1, where the zero values of (b) and (c) are not
the stress-free ones, so those two sweeps (and only those) miss by O(1) next to that wall.  A
stress-free value in place of either zero would make state B converge there.
"""
from __future__ import annotations

import numpy as np
import pytest
import sympy as sp

import oracle
from cssx.model import Config, Model

N = 32


def _sweeps(state, n=N, nphi=32):
    c = Config(nr=n, nth=n, nphi=nphi, kappa_value=0.01, Laplacian=True)
    m = Model(c)
    r, th, ph = sp.symbols("r th ph")
    xi = (r - c.r1) / (c.r2 - c.r1)
    eta = (th - c.th1) / (c.th2 - c.th1)
    up_, dn_ = 1 + sp.cos(sp.pi * xi), 1 - sp.cos(sp.pi * xi)
    gw, gv = (up_, dn_) if state == "A" else (dn_, up_)
    w = r * gw * sp.cos(2 * th) * sp.sin(4 * ph) / 20
    v = r * gv * sp.sin(th) * sp.cos(sp.pi * eta) * sp.cos(4 * ph + sp.Rational(1, 2)) * 7 / 100
    u = sp.cos(3 * r) * sp.cos(sp.pi * eta) * sp.cos(4 * ph) / 10       # du/dth = 0 on theta walls
    sn = sp.sin(th)
    exact = [sp.diff(w, r, th), sp.diff(u, r, ph) / sn, sp.diff(u, r, th), sp.diff(v, r, ph) / sn,
             sp.diff(v, th, ph) / r, sp.diff(w, th, ph) / r]
    R, P, T = np.meshgrid(m.r, m.phi, m.theta, indexing="ij")
    st = np.zeros((8, c.nr, c.nphi, c.nth))
    st[0] = 1.0
    st[4] = 0.1
    for iv, f in ((1, w), (2, v), (3, u)):
        st[iv] = np.broadcast_to(sp.lambdify((r, th, ph), f, "numpy")(R, T, P), R.shape)
    got = oracle.lap_cross(m, st)
    ex = np.stack([np.broadcast_to(sp.lambdify((r, th, ph), e, "numpy")(R, T, P), R.shape) for e in exact])
    return got, ex


def _row_errors(got, ex, k):
    """Relative error of sweep k per row along its own axis (r for k < 4, theta after)."""
    sc = np.abs(ex[k]).max()
    d = np.abs(got[k] - ex[k]) / sc
    return d.max(axis=(1, 2)) if k < 4 else d.max(axis=(0, 1))


@pytest.fixture(scope="module")
def state_a():
    return _sweeps("A")


@pytest.fixture(scope="module")
def state_b():
    return _sweeps("B")


@pytest.mark.parametrize("k", range(6))
def test_stress_free_state_converges_to_the_walls(state_a, k):
    got, ex = state_a
    e = _row_errors(got, ex, k)
    assert e.max() < 1e-4, (k, e[:3], e[-3:])


def test_unwritten_walls_are_zero(state_b):
    got, ex = state_b
    e1 = _row_errors(got, ex, 0)       # (b): X1 != 0 at r2, the wall value used is 0
    e4 = _row_errors(got, ex, 3)       # (c): X4 != 0 at r1, the wall value used is 0
    assert e1[-1] > 0.05 and e1[: N // 2].max() < 1e-4, (e1[:3], e1[-3:])
    assert e4[0] > 0.05 and e4[N // 2:].max() < 1e-4, (e4[:3], e4[-3:])
    for k in (1, 2, 4, 5):
        assert _row_errors(got, ex, k).max() < 1e-4, k
    # the zero value's miss is the compact response to a wrong Neumann datum: it decays away
    # from the wall, row by row
    assert np.all(np.diff(e4[:5]) < 0) and np.all(np.diff(e1[-5:]) > 0)


def test_convergence_order_near_walls():
    """Row errors next to every wall fall off with the closures' order (n = 24 -> 40)."""
    for k in range(6):
        e24 = _row_errors(*_sweeps("A", 24, 32), k)
        e40 = _row_errors(*_sweeps("A", 40, 32), k)
        near24 = max(e24[:3].max(), e24[-3:].max())
        near40 = max(e40[:3].max(), e40[-3:].max())
        assert near40 < near24 or near40 < 2e-6, (k, near24, near40)
