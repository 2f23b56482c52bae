"""Pins the CPU oracle's SLD diffusion (Compute_Diffusion, diffusion.f90:70-1495).

No reference test touches the diffusion, so the oracle is pinned by what the scheme fixes:
  * the fluxes are in conservation form: with zero flux through the walls and periodic phi,
    the volume integral of every diffused field's tendency vanishes (pins the r^2 and
    sin(theta) face weights against the cell volumes (r_h^3-r_l^3)/3 (cos th_l - cos th_h) dphi);
  * the Van Albada form min(1, q- q+/|q-|^2) is minmod: a phi-only field gives
    geom(r,theta) * (delue_{k+1} - delue_k) with delue = q - (minmod(q-,q) + minmod(q+,q))/2,
    written here with numpy, and geom tends to 1/(r sin(theta) dphi);
  * a field linear in the grid index has no interior flux (limiter consistency);
  * sld_coef = 0 gives exactly zero and the tendency is linear in sld_coef;
  * the SLD speed^2 adds |u|^2 and B^2/(4 pi rho) when dstype = 1 (closed-form ratio);
  * the corrections do what diffusion.f90 says they are for: mass diffusion leaves momentum
    rho u and temperature unchanged (:395-428), viscous heating turns the kinetic energy the
    velocity diffusion removes into heat rho T ds (:375-393), Ohmic heating is
    -(1/2) B . curl(dA/dt) / (4 pi rho T) (:430-461, the curl checked against the RHS's B);
  * the SLDiffusion branch of Step_RK4 / Step_RK3_TVD / Step_RK2_TVD adds exactly dt * D to
    the step (step_rk4.f90:366-383, :197-214, :60-77).
The wall ghost rules of Diffusion_R/Th (diffusion.f90:586-711, 1008-1120) are pinned in
test_oracle_sld_walls.py.
"""
import numpy as np
import pytest

import oracle
from cssx.model import Config, Model


def _faces(m):
    c = m.c
    r, dr, th = m.r, m.dr, m.theta
    rl = np.empty(c.nr)
    rh = np.empty(c.nr)
    rl[0] = r[0] - 0.5 * dr[0]
    rl[1:] = 0.5 * (r[1:] + r[:-1])
    rh[:-1] = 0.5 * (r[1:] + r[:-1])
    rh[-1] = r[-1] + 0.5 * dr[-1]
    dth = (th[-1] - th[0]) / (c.nth - 1)
    tl = np.empty(c.nth)
    th_ = np.empty(c.nth)
    tl[0] = th[0] - 0.5 * dth
    tl[1:] = 0.5 * (th[1:] + th[:-1])
    th_[:-1] = 0.5 * (th[1:] + th[:-1])
    th_[-1] = th[-1] + 0.5 * dth
    return rl, rh, tl, th_


def _volumes(m):
    rl, rh, tl, th_ = _faces(m)
    return ((rh ** 3 - rl ** 3) / 3.0)[:, None, None] * (np.cos(tl) - np.cos(th_))[None, None, :] * m.dphi


def _state(m, vel=1e-2, seed=0):
    rng = np.random.default_rng(seed)
    st = m.initial_state()
    shp = st.shape[1:]
    for iv in (1, 2, 3):
        st[iv] = vel * rng.standard_normal(shp)
    st[0] *= 1 + 1e-3 * rng.standard_normal(shp)
    if st.shape[0] == 8:
        st[5:8] *= 1e3
    return st


SMALL = dict(nr=20, nth=18, nphi=16)


@pytest.mark.parametrize("md", [True, False])
def test_conservation_form(md):
    c = Config(**SMALL, sld_coef=0.3, Mass_Diffusion=md, Do_Viscous_Heating=False, Do_Ohmic_Heating=False)
    m = Model(c)
    st = _state(m)
    D = oracle.diffusion(m, st)
    V = _volumes(m)
    conserved = [0, 5, 6, 7] if md else [1, 2, 3, 4, 5, 6, 7]
    for iv in conserved:
        tot = np.sum(V * D[iv])
        mag = np.sum(V * np.abs(D[iv]))
        assert mag > 0, iv
        assert abs(tot) < 1e-12 * mag, (iv, tot, mag)
    # a wrong face weight breaks it: the r^2-weighted radial flux without its r^2 is not conservative
    assert np.abs(np.sum(V * D[4] * (m.r ** 2)[:, None, None])) > 1e-6 * np.sum(V * np.abs(D[4]))


def _minmod(a, b):
    return np.where(a * b > 0, np.sign(a) * np.minimum(np.abs(a), np.abs(b)), 0.0)


def test_phi_flux_is_minmod_and_geometry_tends_to_continuum():
    c = Config(**SMALL, sld_coef=0.7, magnetic=False, dstype=0, Mass_Diffusion=False)
    m = Model(c)
    st = np.zeros((5, c.nr, c.nphi, c.nth))
    st[0] = m.rho_bar[:, None, None]
    f = np.random.default_rng(3).standard_normal(c.nphi) * 1e-2
    st[4] = m.s_bar[:, None, None] + f[None, :, None]
    D = oracle.diffusion(m, st)
    assert np.all(D[:4] == 0.0)
    fm1, fm2, fp1 = np.roll(f, 1), np.roll(f, 2), np.roll(f, -1)
    # face k sits between f[k-1] and f[k]
    qmh, qph, qp3h = fm1 - fm2, f - fm1, fp1 - f
    delue = qph - 0.5 * (_minmod(qmh, qph) + _minmod(qp3h, qph))
    ddel = np.roll(delue, -1) - delue                       # delue_{k+1} - delue_k
    speed = np.sqrt(m.diffusion_speed[1:c.nr + 1])          # sqrt(vcc), r-only here
    shape = c.sld_coef * speed[:, None, None] * ddel[None, :, None]
    kbig = int(np.argmax(np.abs(ddel)))
    geom = D[4][:, kbig, :] / shape[:, kbig, :]
    assert np.abs(D[4] - geom[:, None, :] * shape).max() < 1e-12 * np.abs(D[4]).max()
    cont = 1.0 / (m.r[:, None] * np.sin(m.theta)[None, :] * m.dphi)
    assert np.abs(geom / cont - 1).max() < 3e-3


@pytest.mark.parametrize("axis", ["r", "theta"])
def test_linear_field_has_no_interior_flux(axis):
    c = Config(**SMALL, sld_coef=0.5, magnetic=False, dstype=0, Mass_Diffusion=False)
    m = Model(c)
    st = np.zeros((5, c.nr, c.nphi, c.nth))
    st[0] = m.rho_bar[:, None, None]
    if axis == "r":
        lin = 1e-2 * np.arange(c.nr)[:, None, None]
        sl = (slice(2, c.nr - 2), slice(None), slice(None))
    else:
        lin = 1e-2 * np.arange(c.nth)[None, None, :]
        sl = (slice(None), slice(None), slice(2, c.nth - 2))
    st[4] = m.s_bar[:, None, None] + lin
    D = oracle.diffusion(m, st)
    assert np.abs(D[4][sl]).max() < 1e-14
    assert np.abs(D[4]).max() > 1e-6            # the walls' ghost rows do produce a flux


def test_zero_coefficient_and_linearity():
    c0 = Config(**SMALL, sld_coef=0.0)
    st = _state(Model(c0))
    assert np.all(oracle.diffusion(Model(c0), st) == 0.0)
    D1 = oracle.diffusion(Model(Config(**SMALL, sld_coef=0.25)), st)
    D2 = oracle.diffusion(Model(Config(**SMALL, sld_coef=0.5)), st)
    for iv in range(8):
        assert np.abs(D2[iv] - 2 * D1[iv]).max() <= 1e-12 * np.abs(D2[iv]).max(), iv


def test_speed_includes_flow_and_alfven_speed():
    base = dict(**SMALL, sld_coef=0.7, Mass_Diffusion=False, Do_Viscous_Heating=False, Do_Ohmic_Heating=False)
    c = Config(**base, dstype=0)
    m = Model(c)
    st = np.zeros((8, c.nr, c.nphi, c.nth))
    st[0] = m.rho_bar[:, None, None]
    f = np.random.default_rng(4).standard_normal(c.nphi) * 1e-2
    st[4] = m.s_bar[:, None, None] + f[None, :, None]
    U = np.array([0.3, -0.2, 0.5])
    Bc = np.array([0.4, 0.1, -0.6])
    st[1:4] = U[:, None, None, None]
    B = np.broadcast_to(Bc[:, None, None, None], (3,) + st.shape[1:]).copy()
    D0 = oracle.diffusion(m, st, B=B)
    D1 = oracle.diffusion(Model(Config(**base, dstype=1)), st, B=B)
    sp2 = m.diffusion_speed[1:c.nr + 1][:, None, None]
    expect = np.sqrt((sp2 + U @ U + (Bc @ Bc) / (4 * np.pi * m.rho_bar[:, None, None])) / sp2)
    nz = np.abs(D0[4]) > 1e-12 * np.abs(D0[4]).max()
    assert np.abs(D1[4][nz] / D0[4][nz] / np.broadcast_to(expect, D0[4].shape)[nz] - 1).max() < 1e-12


def test_mass_diffusion_keeps_momentum_and_temperature():
    kw = dict(**SMALL, sld_coef=0.4, Do_Viscous_Heating=False, Do_Ohmic_Heating=False)
    m_on = Model(Config(**kw, Mass_Diffusion=True))
    m_off = Model(Config(**kw, Mass_Diffusion=False))
    st = _state(m_on, vel=3e-2, seed=1)
    st[0] *= 1 + 1e-2 * np.random.default_rng(9).standard_normal(st[0].shape)
    Don = oracle.diffusion(m_on, st)
    Doff = oracle.diffusion(m_off, st)
    rho = st[0]
    assert np.abs(Don[0]).max() > 0 and np.all(Doff[0] == 0)
    for iv in (1, 2, 3):
        # d(rho u) = rho du + u drho must equal rho * (diffusion of u alone)
        lhs = rho * Don[iv] + st[iv] * Don[0]
        assert np.abs(lhs - rho * Doff[iv]).max() < 1e-12 * np.abs(rho * Doff[iv]).max(), iv
    # temperature T ~ rho^(gamma-1) exp(s/Cv): dT/T = (gamma-1) drho/rho + ds/Cv; the mass
    # term alone (on - off) must leave T unchanged
    dT = m_on.gamm1 * Don[0] / rho + (Don[4] - Doff[4]) / m_on.Cv
    assert np.abs(dT).max() < 1e-12 * np.abs(m_on.gamm1 * Don[0] / rho).max()


def test_viscous_heating_converts_kinetic_energy():
    kw = dict(**SMALL, sld_coef=0.4, Mass_Diffusion=False, Do_Ohmic_Heating=False)
    m_on = Model(Config(**kw, Do_Viscous_Heating=True))
    m_off = Model(Config(**kw, Do_Viscous_Heating=False))
    st = _state(m_on, vel=5e-2, seed=2)
    Don = oracle.diffusion(m_on, st)
    Doff = oracle.diffusion(m_off, st)
    rho = st[0]
    T = 1.0 / (st[0] ** (-m_on.gamm1) * np.exp(-st[4] / m_on.Cv) / m_on.invtconst)
    dke = rho * sum(st[iv] * Don[iv] for iv in (1, 2, 3))
    heat = rho * T * (Don[4] - Doff[4])
    assert np.abs(dke).max() > 0
    assert np.abs(dke + heat).max() < 1e-12 * np.abs(dke).max()


def test_ohmic_heating_against_curl_of_da():
    kw = dict(**SMALL, sld_coef=0.4, Mass_Diffusion=False, Do_Viscous_Heating=False)
    m_on = Model(Config(**kw, Do_Ohmic_Heating=True))
    m_off = Model(Config(**kw, Do_Ohmic_Heating=False))
    st = _state(m_on, seed=3)
    _, B = oracle.rhs(m_on, st, want_b=True)
    Don = oracle.diffusion(m_on, st, B=B)
    Doff = oracle.diffusion(m_off, st, B=B)
    for iv in range(8):
        if iv != 4:
            assert np.array_equal(Don[iv], Doff[iv]), iv
    # curl(dA/dt) from the RHS's own curl (B output of a state whose A is dA/dt); the RHS
    # zeroes some B components on wall rows, the Ohmic curl does not: compare off the walls
    sda = st.copy()
    sda[5:8] = Don[5:8]
    _, C = oracle.rhs(m_on, sda, want_b=True)
    rho = st[0]
    T = 1.0 / (st[0] ** (-m_on.gamm1) * np.exp(-st[4] / m_on.Cv) / m_on.invtconst)
    expect = -0.5 / (4 * np.pi) * np.sum(B * C, axis=0) / (rho * T)
    sl = (slice(1, -1), slice(None), slice(1, -1))
    got = Don[4] - Doff[4]
    assert np.abs(got[sl]).max() > 0
    assert np.abs(got[sl] - expect[sl]).max() < 1e-10 * np.abs(expect[sl]).max()


@pytest.mark.parametrize("scheme", ["rk4", "rk3_tvd", "rk2_tvd"])
def test_step_adds_dt_times_diffusion(scheme):
    """The stage states never see D: each scheme's step moves by exactly dt * D(y0) more
    (step_rk4.f90:366-383, :197-214 with :318-323, :60-77 with :135-140)."""
    kw = dict(**SMALL)
    m_sld = Model(Config(**kw, sld_coef=0.3))
    m_0 = Model(Config(**kw, sld_coef=0.0))
    st = _state(m_sld, seed=5)
    dt = 2e-4
    D = oracle.diffusion(m_sld, st)
    y1 = oracle.rk4(m_sld, st, dt, 1, scheme)
    y0 = oracle.rk4(m_0, st, dt, 1, scheme)
    sl = (slice(None), slice(1, -1), slice(None), slice(1, -1))       # off the rows Bound sets
    diff = (y1 - y0)[sl]
    want = dt * D[sl]
    for iv in range(8):
        tol = 1e-9 * np.abs(want[iv]).max() + 8e-16 * np.abs(st[iv]).max()
        assert np.abs(diff[iv] - want[iv]).max() <= tol, iv
