"""GPU parity of the explicit-diffusion RHS (Laplacian / SLaplacian: rhs_mpi.f90:844-1474,
:2070-2366, VectorCurl derivatives.f90:1894-2104) and its RK4 stepping and CFL limit against the
CPU oracle.

The manufactured state of tests/mms_visc.py at small uneven shapes, diffusion strong enough
(kappa 1e-2) that the explicit terms are a large part of every tendency.  Both the full RHS and
the explicit contribution alone (RHS with - RHS without) are compared, so the parity of the
new terms is not hidden behind the ideal ones.  Tolerances: FP64 1e-9 of each field's max norm
(the second derivatives carry the rounding of |f| times dx^-2, ~1e3 above the first
derivatives'; the contribution 1e-8: it is a difference); FP32 near the background as the plain
RHS tests.
"""
import dataclasses

import numpy as np
import pytest
import torch

import mms_visc
import oracle
from cssx.model import Config, Model

pytestmark = pytest.mark.gpu

SHAPES = [(24, 40, 32), (40, 36, 72)]      # (nr, nth, nphi)


def _rel_fields(a, b):
    return np.array([np.abs(a[iv] - b[iv]).max() / max(np.abs(b[iv]).max(), 1e-300) for iv in range(a.shape[0])])


def _pair(c, st):
    """(GPU rhs, oracle rhs) of state st for config c, FP64."""
    from cssx.mhd import Solver
    m = Model(c)
    sol = Solver(m)
    got = sol.rhs(torch.as_tensor(st, device="cuda")).cpu().numpy()
    sol.close()
    return got, oracle.rhs(m, st)


def _check(c, tol=1e-9, tol_contrib=1e-8):
    st = mms_visc.evaluate(Model(c))[0][: c.nv]
    print(f"laplacian rhs {c.nr}x{c.nth}x{c.nphi} mag={c.magnetic} lap={c.Laplacian} slap={c.SLaplacian}",
          flush=True)
    got, ref = _pair(c, st)
    e = _rel_fields(got, ref)
    assert e.max() < tol, e
    plain = dataclasses.replace(c, Laplacian=False, SLaplacian=False)
    got0, ref0 = _pair(plain, st)
    # fields where the explicit terms are a real part of the tendency (SLaplacian reaches the
    # momentum only through a wall value: there the difference is cancellation-dominated)
    sel = [iv for iv in range(c.nv) if np.abs(ref[iv] - ref0[iv]).max() > 1e-3 * np.abs(ref[iv]).max()]
    assert sel
    ec = _rel_fields((got - got0)[sel], (ref - ref0)[sel])
    assert ec.max() < tol_contrib, (sel, ec)
    return got, ref


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("magnetic", [True, False])
def test_laplacian_rhs_matches_oracle(shape, magnetic):
    nr, nth, nphi = shape
    _check(Config(nr=nr, nth=nth, nphi=nphi, omega0=0.3, damping_time=50.0, kappa_value=0.01, g0=0.5,
                  Laplacian=True, magnetic=magnetic, lumr1=0.3, lumr2=0.2))


@pytest.mark.parametrize("walls", [dict(Bottom_Stress_Free=False), dict(Top_Stress_Free=False),
                                   dict(Top_Constant_Temp=False, Top_Constant_Flux=True),
                                   dict(Theta_Normal_Field=False, Theta_Stress_Free=False)])
def test_laplacian_wall_variants(walls):
    _check(Config(nr=24, nth=40, nphi=32, kappa_value=0.01, Laplacian=True, lumr1=0.3, lumr2=0.2, **walls))


@pytest.mark.parametrize("flux", [dict(), dict(Top_Constant_Temp=False, Top_Constant_Flux=True)])
def test_slaplacian_rhs_matches_oracle(flux):
    _check(Config(nr=40, nth=36, nphi=72, kappa_value=0.01, g0=1.0, SLaplacian=True, lumr1=0.3, lumr2=0.2,
                  **flux))


def _near_background(c):
    m = Model(c)
    st = m.initial_state()
    rng = np.random.default_rng(5)
    R, P, T = np.meshgrid(m.r, m.phi, m.theta, indexing="ij")
    st[1] += 1e-2 * np.sin(3 * R) * np.sin(2 * T) * np.cos(4 * P)
    st[2] += 1e-2 * np.cos(2 * R) * np.sin(T) * np.sin(4 * P + 0.3)
    st[3] += 1e-2 * np.sin(5 * R) * np.sin(3 * T) * np.cos(4 * P + 1.0)
    st[4] *= 1 + 1e-4 * rng.standard_normal(st[4].shape)
    return m, st


def test_laplacian_f32_near_background():
    """FP32 fluctuation form: the explicit terms (including the background's second radial
    entropy derivative) against the FP64 oracle, at the precision of the plain FP32 RHS."""
    from cssx.mhd import Solver
    c = Config(nr=40, nth=36, nphi=72, kappa_value=1e-3, Laplacian=True, lumr1=0.3, lumr2=0.2)
    m, st = _near_background(c)
    print("laplacian f32", flush=True)
    ref = oracle.rhs(m, st)
    sol = Solver(m)
    x = torch.as_tensor(st, device="cuda")
    got32 = sol.rhs(sol.to_f32(x)).double().cpu().numpy()
    bg = sol.background("cpu").numpy()
    sol.close()
    e32 = _rel_fields(got32, ref)
    # as the plain FP32 test: the density tendency within 20x of what rounding the input to the
    # FP32 fluctuation form does to the exact RHS, u_r within FP32 rounding of gravity; the
    # other fields add the FP32 rounding of the second-derivative sweeps (eps32 |f| dx^-2)
    xr = (st - bg).astype(np.float32).astype(np.float64) + bg
    cond = _rel_fields(oracle.rhs(m, xr), ref)
    print("laplacian f32 rel err per field", e32, "input-rounding sensitivity", cond)
    eps32 = float(np.finfo(np.float32).eps)
    assert e32[0] < 20 * cond[0] + 1e-6, (e32, cond)
    assert e32[3] < 64 * eps32 * m.gravity.max() / np.abs(ref[3]).max(), e32
    others = [1, 2, 4, 5, 6, 7]
    assert np.all(e32[others] < 20 * cond[others] + 2e-5), (e32, cond)


@pytest.mark.parametrize("sld", [True, False])
def test_laplacian_rk4_matches_oracle(sld):
    """SLDiffusion with sld_coef 0: classical weights; without SLDiffusion the reference's k1
    weight 1 (step_rk4.f90:387)."""
    from cssx.mhd import Solver
    c = Config(nr=24, nth=40, nphi=32, kappa_value=1e-3, Laplacian=True, SLDiffusion=sld)
    m, st = _near_background(c)
    st = oracle.bound(m, st)
    print(f"laplacian rk4 sld={sld}", flush=True)
    sol = Solver(m)
    y = torch.as_tensor(st, device="cuda").contiguous()
    dt = sol.cfl(y)
    ref = oracle.rk4(m, st, dt, 2)
    sol.rk4(y, dt, 2, "f64")
    # compare the change over the two steps (the state itself is dominated by the background)
    e = _rel_fields(y.cpu().numpy() - st, ref - st)
    assert e.max() < 1e-9, e
    y32 = torch.as_tensor(st, device="cuda").contiguous()
    sol.rk4(y32, dt, 2, "f32")
    sol.close()
    d0 = np.abs(ref - st).max()
    assert np.abs(y32.cpu().numpy() - ref).max() < 1e-3 * d0


def test_laplacian_cfl_limit():
    """Compute_Step_Size with Laplacian: the diffusive limit (rhs_mpi.f90:194-198) takes over
    once the diffusivities are large."""
    from cssx.mhd import Solver
    c = Config(nr=24, nth=40, nphi=32, kappa_value=1.0, Laplacian=True)
    m = Model(c)
    st = m.initial_state()
    sol = Solver(m)
    dt = sol.cfl(torch.as_tensor(st, device="cuda"))
    sol.close()
    dx = min(m.dr.min(), c.r1 * min(min(np.sin(m.theta[0]), np.sin(m.theta[-1])) * m.dphi, m.dth))
    dtd = dx ** 2 / max(m.ks.max(), (m.mur / m.rho_bar).max(), m.etar.max())
    assert dt == pytest.approx(dtd * c.cfl_safety, rel=1e-12)
