"""Pins the oracle's explicit-diffusion branches (Laplacian / SLaplacian, rhs_mpi.f90:844-1474,
:2070-2366, VectorCurl derivatives.f90:1894-2104) against the continuous equations.

The Laplacian contribution is RHS(Laplacian) - RHS(plain) of one manufactured state
(tests/mms_visc.py): in the interior it must converge to div(tau)/rho, the thermal conduction and
viscous dissipation over rho T, and -eta curl B with its Ohmic heating, plus four closed-form
departures of the reference from those formulas (each cited in mms_visc).  Each departure is
>100x the residual, so leaving one out (or adding one the oracle does not have) fails.
"""
import dataclasses

import numpy as np
import pytest

import mms_visc
import oracle
from cssx.model import Config, Model


def _contribution(n, **kw):
    c = Config(nr=n, nth=n, nphi=n, damping_time=1e9, kappa_value=0.01, **kw)
    m = Model(c)
    st, exp, devs, cond = mms_visc.evaluate(m)
    plain = Model(dataclasses.replace(c, Laplacian=False, SLaplacian=False))
    lap = oracle.rhs(m, st) - oracle.rhs(plain, st)
    return lap, exp, devs, cond


def _interior(n):
    b = n // 4
    return (slice(b, n - b), slice(None), slice(b, n - b))


@pytest.fixture(scope="module")
def lap48():
    return _contribution(48, Laplacian=True)


def _errors(n, lap, exp, devs):
    sl = _interior(n)
    tot = exp + sum(devs.values())
    return [np.abs(lap[iv][sl] - tot[iv][sl]).max() / np.abs(tot[iv][sl]).max() for iv in range(1, 8)]


def test_laplacian_matches_continuous_terms(lap48):
    lap, exp, devs, _ = lap48
    err = _errors(48, lap, exp, devs)
    # momentum and entropy: 1e-5 of their scale; the resistive terms (curl of curl of A, and
    # the walls of Compute_B) 1e-2
    assert max(err[:4]) < 1e-5, err
    assert max(err[4:]) < 1e-2, err


def test_laplacian_convergence():
    lap, exp, devs, _ = _contribution(32, Laplacian=True)
    e32 = _errors(32, lap, exp, devs)
    e48 = _errors(48, *_contribution(48, Laplacian=True)[:3])
    order = [np.log(a / b) / np.log(48 / 32) for a, b in zip(e32, e48)]
    assert min(order) > 5.0, (e32, e48, order)


def test_each_departure_matters(lap48):
    lap, exp, devs, _ = lap48
    sl = _interior(48)
    tot = exp + sum(devs.values())
    resid = max(np.abs(lap[iv][sl] - tot[iv][sl]).max() for iv in (1, 2, 3, 4))
    for name, d in devs.items():
        assert np.abs(d[(slice(None),) + sl]).max() > 100 * resid, name
    # and the viscous force itself, per component
    for iv in (1, 2, 3):
        assert np.abs(exp[iv][sl]).max() > 1000 * resid, iv


def test_slaplacian_is_thermal_conduction_only():
    n = 48
    lap, exp, devs, cond = _contribution(n, SLaplacian=True)
    sl = _interior(n)
    want = cond + devs["s_th"][4]
    assert np.abs(lap[4][sl] - want[sl]).max() < 1e-5 * np.abs(want[sl]).max()
    for iv in (1, 2, 3, 5, 6, 7):
        # momentum and induction see the thermal branch only through the constant-flux wall
        # value of ds/dr, which the interior does not feel
        assert np.abs(lap[iv][sl]).max() < 1e-6 * np.abs(exp[iv][sl]).max(), iv


def test_density_closure_absorbs_viscous_terms_at_walls():
    """The explicit terms enter the momentum tendency before Density_Terms_Theta/Radial, whose
    wall derivatives then zero du/dt on the radial walls and dw/dt on the theta walls
    (rhs_mpi.f90:1553-1564, :1750-1773) - viscous part included."""
    n = 32
    c = Config(nr=n, nth=n, nphi=n, damping_time=1e9, kappa_value=0.01, Laplacian=True)
    m = Model(c)
    st = mms_visc.evaluate(m)[0]
    dv = oracle.rhs(m, st)
    du, dw = dv[3], dv[1]
    assert np.abs(du[0]).max() < 1e-10 * np.abs(du).max()
    assert np.abs(du[-1]).max() < 1e-10 * np.abs(du).max()
    assert np.abs(dw[:, :, 0]).max() < 1e-10 * np.abs(dw).max()
    assert np.abs(dw[:, :, -1]).max() < 1e-10 * np.abs(dw).max()
