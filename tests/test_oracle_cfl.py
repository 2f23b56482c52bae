"""Closed-form pins of the oracle's Compute_Step_Size (rhs_mpi.f90:92-275).

On a uniform radial grid (``stretch=0``: every ``dr`` equals h) and uniform states, every
quantity Compute_Step_Size forms has a closed form:

* the crossing-time minimum sits at the smallest cell, r1 sin(th1) dphi when the grid is
  chosen so that the phi width of the first row is the narrowest; then
  dt = cfl * r1 sin(th1) dphi / sqrt(|u|^2 + c^2) with c^2 = (gamma-1) Cp xi_cs^2 T;
* the Alfven limit uses the textbook Alfven speed B/sqrt(4 pi rho) of an analytic uniform
  field B0 z (vector potential A_phi = B0 r sin(th)/2);
* the conservation totals are products of power sums of the radii and sums of sin^k over an
  arithmetic progression of colatitudes, each of which has a closed form;
* the explicit-diffusion limit is x^2 / (kappa rho_bar(r1) T_bar(r1)) for the polytrope.

A dropped factor (xi_cs, gamma-1, 4 pi, r or sin(th) in the cell width, a wrong power of r or
sin(th) in a total) moves the result by far more than the tolerances.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from cssx.model import Config, Model


def _model(**kw):
    base = dict(nr=8, nth=16, nphi=128, stretch=0.0)
    base.update(kw)
    return Model(Config(**base))


def _uniform(m, rho0=2.0, s0=0.3, w=0.0, v=0.0, u=0.0):
    c = m.c
    st = np.zeros((c.nv, c.nr, c.nphi, c.nth))
    st[0] = rho0
    st[1], st[2], st[3] = w, v, u
    st[4] = s0
    return st


def _temperature(m, rho0, s0):
    return m.invtconst * rho0 ** m.gamm1 * np.exp(s0 / m.Cv)


def _dmin(m):
    c = m.c
    dphi = (c.phi2 - c.phi1) / c.nphi
    d = c.r1 * np.sin(c.th1) * dphi
    # the grid is chosen so that the first row's phi width is the narrowest cell width
    assert d < (c.r2 - c.r1) / (c.nr - 1) and d < c.r1 * (c.th2 - c.th1) / (c.nth - 1)
    return d


def test_uniform_grid_has_equal_cells():
    m = _model()
    h = (m.c.r2 - m.c.r1) / (m.c.nr - 1)
    assert np.allclose(m.dr, h, rtol=1e-12, atol=0)


@pytest.mark.parametrize("xi", [1.0, 0.5])
def test_sound_crossing_time_closed_form(xi):
    m = _model(xi_cs=xi)
    rho0, s0 = 2.0, 0.3
    got = oracle.step_size(m, _uniform(m, rho0, s0))
    c2 = m.gamm1 * m.c.Cp * xi * xi * _temperature(m, rho0, s0)
    ref = m.c.cfl_safety * _dmin(m) / np.sqrt(c2)
    assert abs(got["dt"] - ref) < 1e-13 * ref
    assert got["dtmaxb"] == got["dtmaxc"]          # A = 0: B = 0, no Alfven term
    assert got["dtmaxd"] == 1e10


@pytest.mark.parametrize("comp", ["w", "v", "u"])
def test_flow_speed_adds_in_quadrature(comp):
    m = _model()
    rho0, s0, U = 1.5, -0.2, 0.7
    got = oracle.step_size(m, _uniform(m, rho0, s0, **{comp: U}))
    c2 = m.gamm1 * m.c.Cp * _temperature(m, rho0, s0)
    ref = m.c.cfl_safety * _dmin(m) / np.sqrt(U * U + c2)
    assert abs(got["dt"] - ref) < 1e-13 * ref


def test_cfl_safety_scales_linearly():
    m = _model()
    st = _uniform(m)
    a = oracle.step_size(m, st, 0.4)["dt"]
    b = oracle.step_size(m, st, 0.1)["dt"]
    assert abs(a - 4 * b) < 1e-14 * a


def test_alfven_limit_of_a_uniform_field():
    """B = B0 z from A_phi = B0 r sin(th)/2; Theta_Perf_Conductor zeroes B_theta on the theta
    walls (derivatives.f90:1840-1883) and the radial-field flags are off, so |B|^2 = B0^2
    inside and B0^2 cos^2(th) on the theta walls.  dtmaxb is the minimum over cells of
    dxmin / sqrt(c^2 + |B|^2/(4 pi rho0))."""
    m = _model(nr=12, nth=24, nphi=96, Top_Radial_Field=False, Bottom_Radial_Field=False,
               Theta_Normal_Field=False, Theta_Perf_Conductor=True)
    c = m.c
    rho0, s0, B0 = 1.0, 0.0, 3.0
    st = _uniform(m, rho0, s0)
    r = m.r[:, None, None]
    th = m.theta[None, None, :]
    st[6] = 0.5 * B0 * r * np.sin(th) * np.ones((1, c.nphi, 1))
    got = oracle.step_size(m, st)
    c2 = m.gamm1 * c.Cp * _temperature(m, rho0, s0)
    dphi = (c.phi2 - c.phi1) / c.nphi
    h = m.dr[0]
    rr = m.r[:, None]
    sn = np.sin(m.theta)[None, :]
    dx = np.minimum(h, rr * np.minimum(m.dth, sn * dphi))
    b2 = np.full((c.nr, c.nth), B0 * B0)
    b2[:, 0] = (B0 * np.cos(m.theta[0])) ** 2
    b2[:, -1] = (B0 * np.cos(m.theta[-1])) ** 2
    va2 = b2 / (4 * np.pi * rho0)
    ref_b = (dx / np.sqrt(c2 + va2)).min()
    assert va2.max() > 0.3 * c2                    # the Alfven term moves the answer
    assert abs(got["dtmaxb"] - ref_b) < 2e-6 * ref_b    # compact derivative of sin^2: O(dth^6)
    assert abs(got["dtmaxc"] - _dmin(m) / np.sqrt(c2)) < 1e-13 * got["dtmaxc"]
    assert abs(got["dt"] - c.cfl_safety * ref_b) < 2e-6 * got["dt"]


# closed-form sums over arithmetic progressions
def _power_sum(a, h, n, k):
    """sum_{i<n} (a + i h)^k for k <= 3 from the Faulhaber sums of i^j."""
    s = [n, n * (n - 1) / 2, (n - 1) * n * (2 * n - 1) / 6, (n * (n - 1) / 2) ** 2]
    from math import comb
    return sum(comb(k, j) * a ** (k - j) * h ** j * s[j] for j in range(k + 1))


def _sin_sum(a, d, n, k=1):
    """sum_{j<n} sin(a + j d) (k=1), sin^2 (k=2) or sin^3 (k=3), closed forms from
    sum sin(x + j y) = sin(n y/2) sin(x + (n-1) y/2) / sin(y/2) and the power reductions
    sin^2 = (1 - cos 2x)/2, sin^3 = (3 sin x - sin 3x)/4."""
    def ssin(x, y):
        return np.sin(n * y / 2) * np.sin(x + (n - 1) * y / 2) / np.sin(y / 2)

    def scos(x, y):
        return np.sin(n * y / 2) * np.cos(x + (n - 1) * y / 2) / np.sin(y / 2)
    if k == 1:
        return ssin(a, d)
    if k == 2:
        return 0.5 * n - 0.5 * scos(2 * a, 2 * d)
    return 0.75 * ssin(a, d) - 0.25 * ssin(3 * a, 3 * d)


def test_conservation_totals_closed_form():
    m = _model(nr=10, nth=12, nphi=16)
    c = m.c
    rho0, s0, V = 1.7, 0.25, 0.6
    got = oracle.step_size(m, _uniform(m, rho0, s0, v=V))
    h = m.dr[0]
    dphi = (c.phi2 - c.phi1) / c.nphi
    R2, R3 = _power_sum(c.r1, h, c.nr, 2), _power_sum(c.r1, h, c.nr, 3)
    S1, S2, S3 = (_sin_sum(c.th1, m.dth, c.nth, k) for k in (1, 2, 3))
    ang = m.dth * c.nphi * dphi * h
    T = _temperature(m, rho0, s0)
    ref = {"m": rho0 * R2 * S1 * ang, "mrsint": rho0 * R3 * S3 * ang, "L": rho0 * V * R3 * S2 * ang,
           "E": rho0 * (m.Cv * T + 0.5 * V * V) * R2 * S1 * ang}
    for k, v in ref.items():
        assert abs(got[k] - v) < 1e-12 * abs(v), (k, got[k], v)


def test_rotation_enters_angular_momentum():
    """v -> v + r sin(th) Omega0: at rest in a rotating frame L = rho0 Omega0 sum r^4 sin^3."""
    m = _model(nr=10, nth=12, nphi=16, omega0=0.3)
    c = m.c
    rho0 = 1.2
    got = oracle.step_size(m, _uniform(m, rho0, 0.0))
    h = m.dr[0]
    dphi = (c.phi2 - c.phi1) / c.nphi
    R4 = sum((c.r1 + i * h) ** 4 for i in range(c.nr))
    ref = rho0 * c.omega0 * R4 * _sin_sum(c.th1, m.dth, c.nth, 3) * m.dth * c.nphi * dphi * h
    assert abs(got["L"] - ref) < 1e-12 * ref


def test_explicit_diffusion_limit():
    m = _model(Laplacian=True, kappa_value=2e-3)
    c = m.c
    got = oracle.step_size(m, _uniform(m))
    dphi = (c.phi2 - c.phi1) / c.nphi
    x = min(m.dr[0], c.r1 * min(np.sin(c.th1) * dphi, m.dth))
    # polytrope: ks = kappa rho_bar T_bar peaks at the bottom, above mur/rho_bar = kappa
    ks1 = c.kappa_value * m.rho_bar[0] * m.T_bar[0]
    assert ks1 > c.kappa_value
    assert abs(got["dtmaxd"] - x * x / ks1) < 1e-13 * got["dtmaxd"]
