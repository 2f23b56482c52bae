"""Pins the CPU oracle's MHD right-hand side, Bound and RK4 stepping.

The reference's own tests hold nothing for RHS/RK4 (no fixture or known answer pins them), so
the oracle is pinned by what the mathematics fixes:
  * manufactured solution (tests/mms.py): the interior RHS converges to the continuous
    equations evaluated with sympy's spherical operators, and every physical term is far
    above the residual, so a dropped term or a flipped sign fails;
  * hydrostatic balance of the isentropic polytrope at rest;
  * the wall conditions the density closure is built to enforce: du/dt = 0 on the radial
    walls, dw/dt = 0 on the theta walls (rhs_mpi.f90:1553-1564, :1750-1773);
  * Bound zeroes exactly the documented wall values;
  * RK4 is fourth order in time (Richardson ratio ~17 over dt, dt/2, dt/4).
"""
import numpy as np
import pytest

import mms
import oracle
from cssx.model import Config, Model


@pytest.fixture(scope="module")
def mms64():
    c = Config(nr=64, nth=64, nphi=64, omega0=0.3, damping_time=1e9)
    m = Model(c)
    st, ex, terms, Bex = mms.evaluate(m)
    dv, B = oracle.rhs(m, st, want_b=True)
    return m, st, ex, terms, Bex, dv, B


def _interior(n):
    b = n // 4
    return (slice(b, n - b), slice(None), slice(b, n - b))


def test_mms_interior_matches_continuous_equations(mms64):
    m, st, ex, terms, Bex, dv, B = mms64
    sl = _interior(64)
    smallest = min(np.abs(v[(slice(None),) + sl]).max() for v in terms.values())
    for iv in range(8):
        err = np.abs(dv[iv][sl] - ex[iv][sl]).max()
        scale = np.abs(ex[iv][sl]).max()
        if iv in (1, 2, 3):
            # momentum: the residual is 100x below the smallest physical term
            assert err < 0.01 * smallest, (iv, err, smallest)
        assert err < 1e-4 * max(scale, 1.0), (iv, err, scale)
    assert np.abs(B[(slice(None),) + sl] - Bex[(slice(None),) + sl]).max() < 1e-6


def test_mms_every_term_matters(mms64):
    """Each momentum term alone is >100x the oracle's residual: removing it would fail."""
    m, st, ex, terms, Bex, dv, B = mms64
    sl = _interior(64)
    resid = max(np.abs(dv[iv][sl] - ex[iv][sl]).max() for iv in (1, 2, 3))
    for name, v in terms.items():
        assert np.abs(v[(slice(None),) + sl]).max() > 100 * resid, name


def test_mms_convergence_order():
    errs = []
    for n in (32, 48):
        c = Config(nr=n, nth=n, nphi=n, omega0=0.3, damping_time=1e9)
        m = Model(c)
        st, ex, terms, Bex = mms.evaluate(m)
        dv = oracle.rhs(m, st)
        sl = _interior(n)
        errs.append(max(np.abs(dv[iv][sl] - ex[iv][sl]).max() / np.abs(ex[iv][sl]).max() for iv in range(8)))
    order = np.log(errs[0] / errs[1]) / np.log(48 / 32)
    assert order > 5.0, (errs, order)


def test_hydrostatic_rest_state():
    c = Config(nr=48, nth=24, nphi=16, magnetic=False, damping_time=1e9)
    m = Model(c)
    st = np.zeros((5, c.nr, c.nphi, c.nth))
    st[0] = m.rho_bar[:, None, None]
    st[4] = m.s_bar[:, None, None]
    dv = oracle.rhs(m, st)
    g = m.gravity.max()
    assert np.abs(dv[3]).max() < 1e-5 * g           # gravity balanced by the pressure gradient
    assert np.abs(dv[[1, 2]]).max() < 1e-12 * g
    assert np.abs(dv[[0, 4]]).max() < 1e-12


def test_density_closure_zeroes_normal_momentum_at_walls(mms64):
    m, st, ex, terms, Bex, dv, B = mms64
    du, dw = dv[3], dv[1]
    assert np.abs(du[0]).max() < 1e-10 * np.abs(du).max()
    assert np.abs(du[-1]).max() < 1e-10 * np.abs(du).max()
    assert np.abs(dw[:, :, 0]).max() < 1e-10 * np.abs(dw).max()
    assert np.abs(dw[:, :, -1]).max() < 1e-10 * np.abs(dw).max()
    # the interior is not zeroed
    assert np.abs(du[10]).max() > 1e-3 * np.abs(du).max()


def test_bound_sets_documented_walls():
    c = Config(nr=16, nth=12, nphi=8)
    m = Model(c)
    rng = np.random.default_rng(5)
    st = rng.standard_normal((8, 16, 8, 12))
    b = oracle.bound(m, st)
    # w at theta walls, u_r at both radial walls, A_theta at theta walls, A_r at radial walls
    assert np.all(b[1][:, :, [0, -1]] == 0)
    assert np.all(b[3][[0, -1]] == 0)
    assert np.all(b[5][:, :, [0, -1]] == 0)
    assert np.all(b[7][[0, -1]] == 0)
    changed = (b != st)
    expect = np.zeros_like(changed)
    expect[1][:, :, [0, -1]] = True
    expect[3][[0, -1]] = True
    expect[5][:, :, [0, -1]] = True
    expect[7][[0, -1]] = True
    assert np.array_equal(changed, expect)


def test_rk4_fourth_order_in_time():
    c = Config(nr=20, nth=16, nphi=16, omega0=0.3, damping_time=1e9)
    m = Model(c)
    st, ex, terms, Bex = mms.evaluate(m, amp=0.2)
    st = oracle.bound(m, st)
    dt = 4e-4
    y1 = oracle.rk4(m, st, dt, 2)
    y2 = oracle.rk4(m, st, dt / 2, 4)
    y4 = oracle.rk4(m, st, dt / 4, 8)
    ratio = np.abs(y1 - y4).max() / np.abs(y2 - y4).max()
    assert 14 < ratio < 20, ratio


@pytest.mark.parametrize("scheme,expect", [("rk3_tvd", 9.0), ("rk2_tvd", 5.0)])
def test_tvd_schemes_order_in_time(scheme, expect):
    """Step_RK3_TVD / Step_RK2_TVD: Richardson ratio (1-4^-p)/(2^-p-4^-p) = 9 (p=3), 5 (p=2)."""
    c = Config(nr=20, nth=16, nphi=16, omega0=0.3, damping_time=1e9)
    m = Model(c)
    st, ex, terms, Bex = mms.evaluate(m, amp=0.2)
    st = oracle.bound(m, st)
    dt = 4e-4
    y1 = oracle.rk4(m, st, dt, 2, scheme)
    y2 = oracle.rk4(m, st, dt / 2, 4, scheme)
    y4 = oracle.rk4(m, st, dt / 4, 8, scheme)
    ratio = np.abs(y1 - y4).max() / np.abs(y2 - y4).max()
    assert abs(ratio / expect - 1) < 0.15, ratio
