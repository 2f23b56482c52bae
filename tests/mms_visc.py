"""Manufactured solution for the explicit-diffusion (Laplacian / SLaplacian) terms of the RHS.

The fields of tests/mms.py's manufactured state, with the polytrope's diffusion profiles as
sympy expressions (mhdsph_mpi.f90:349-354): mu = sld_pr kappa rho_bar(r),
ks = kappa rho_bar T_bar (rho_bar/rho_bar(r2))^g0, eta = ks.  The continuous terms are built with
sympy.vector's spherical operators and the textbook spherical strain components:
  momentum   div(tau)/rho,  tau = mu (grad u + grad u^T - 2/3 div(u) I)
             = [mu (4/3 grad div u - curl curl u) + mu'(r) tau0.r_hat] / rho
  entropy    [div(ks grad s) + Phi] / (rho T),
             Phi = mu [4/3 (err^2+ett^2+epp^2-err ett-ett epp-epp err) + g_rt^2 + g_rp^2 + g_tp^2]
  induction  dA/dt += -eta curl B,  ds/dt += (4 pi / c^2) eta |curl B|^2 / (rho T)
plus, separately, the reference's own departures from those formulas, each a closed form:
  dev_du     the (1/3) cot(theta) (1/r) du_theta/dr term of Radial_Terms carries no mu/rho
             factor (rhs_mpi.f90:1322): (1 - mu/rho) (1/3) cot (1/r) d(u_theta)/dr
  dev_s_th   the theta second-derivative thermal term scales with 1/r^3, not 1/r^2
             (rhs_mpi.f90:970-973): ks s_thth (1/r^3 - 1/r^2) / (rho T)
  dev_s_tt   the theta-theta / phi-phi heating factor reads e_tt - e_pp + 2 u_r/r where the
             dissipation has e_tt - e_pp (rhs_mpi.f90:1131): (8/3) mu (u_r/r) e_tt / (rho T)
  dev_s_rp   the r-phi heating term is added a second time with du_r/dr in place of du_phi/dr
             (rhs_mpi.f90:1418): mu (du_r/dr + du_r/dphi/(r sin) - u_phi/r)^2 / (rho T)
"""
from __future__ import annotations

import numpy as np
import sympy as sp
from sympy.vector import CoordSys3D, curl, divergence, gradient

C_LIGHT = 2.99792458e10


def default_fields(model, r, th, ph):
    """(rho, s, u_r, u_theta, u_phi, A_r, A_theta, A_phi) of the interior manufactured state."""
    R = sp.Rational
    rho = 1 + R(3, 10) * sp.cos(5 * r) * (2 + sp.sin(2 * th + R(3, 10)) * sp.cos(4 * ph))
    s = -R(1, 2) + R(1, 20) * sp.sin(4 * r) * sp.cos(3 * th) * sp.sin(4 * ph + R(1, 5))
    ur = R(1, 10) * sp.cos(3 * r) * sp.sin(th) * sp.cos(4 * ph)
    ut = R(1, 20) * sp.sin(2 * r) * sp.cos(2 * th) * sp.sin(4 * ph)
    up = R(7, 100) * sp.cos(r) * sp.sin(3 * th) * sp.cos(4 * ph + R(1, 2))
    Ar = R(1, 10) * sp.sin(r) * sp.sin(th) * sp.cos(4 * ph + 1)
    At = R(1, 10) * sp.sin(2 * r) * sp.cos(th) * sp.cos(4 * ph)
    Ap = R(2, 10) * sp.cos(3 * r) * sp.sin(2 * th) * sp.sin(4 * ph)
    return rho, s, ur, ut, up, Ar, At, Ap


def build(model, fields=default_fields):
    c = model.c
    S = CoordSys3D("S", transformation="spherical", variable_names=("r", "theta", "phi"))
    r, th, ph = S.r, S.theta, S.phi
    R = sp.Rational
    rho, s, ur, ut, up, Ar, At, Ap = fields(model, r, th, ph)
    # polytrope profiles
    gamm1 = c.gamma - 1.0
    zeta = 1 + model.alp * (c.r2 / r - 1)
    rho_bar = c.rho0 * zeta ** (1.0 / gamm1)
    T_bar = c.T0 * zeta
    rho_bar2 = c.rho0
    mu = c.sld_pr * c.kappa_value * rho_bar
    ks = c.kappa_value * rho_bar * T_bar * (rho_bar / rho_bar2) ** c.g0
    eta = ks
    Cv = c.Cp / c.gamma
    T = (1.0 / Cv / gamm1) * rho ** gamm1 * sp.exp(s / Cv)
    sn, cs = sp.sin(th), sp.cos(th)

    U = ur * S.i + ut * S.j + up * S.k
    divu = divergence(U)
    visc0 = R(4, 3) * gradient(divu) - curl(curl(U))
    err = sp.diff(ur, r)
    g_rt = r * sp.diff(ut / r, r) + sp.diff(ur, th) / r
    g_rp = r * sp.diff(up / r, r) + sp.diff(ur, ph) / (r * sn)
    g_tp = sn / r * sp.diff(up / sn, th) + sp.diff(ut, ph) / (r * sn)
    tau_r = (2 * err - R(2, 3) * divu) * S.i + g_rt * S.j + g_rp * S.k
    force = (mu * visc0 + sp.diff(mu, r) * tau_r) / rho
    ett = sp.diff(ut, th) / r + ur / r
    epp = sp.diff(up, ph) / (r * sn) + ur / r + ut * cs / (sn * r)
    Phi = mu * (R(4, 3) * (err ** 2 + ett ** 2 + epp ** 2 - err * ett - ett * epp - epp * err)
                + g_rt ** 2 + g_rp ** 2 + g_tp ** 2)
    cond = divergence(ks * gradient(s))
    rT = rho * T
    ds = (cond + Phi) / rT
    A = Ar * S.i + At * S.j + Ap * S.k
    B = curl(A)
    J = curl(B)
    dA = -eta * J
    ds_ohm = (4 * sp.pi / C_LIGHT ** 2) * eta * J.dot(J) / rT
    dev = {
        "du": (1 - mu / rho) * R(1, 3) * (cs / sn) / r * sp.diff(ut, r),
        "s_th": ks * sp.diff(s, th, 2) * (1 / r ** 3 - 1 / r ** 2) / rT,
        "s_tt": R(8, 3) * mu * (ur / r) * ett / rT,
        "s_rp": mu * (err + sp.diff(ur, ph) / (r * sn) - up / r) ** 2 / rT,
    }
    fields = [rho, ut, up, ur, s, At, Ap, Ar]
    args = (r, th, ph)
    out = {
        "fields": sp.lambdify(args, fields, "numpy"),
        # slots (w, v, u) = (theta, phi, r) components
        "force": sp.lambdify(args, [force.dot(S.j), force.dot(S.k), force.dot(S.i)], "numpy"),
        "ds": sp.lambdify(args, ds, "numpy"),
        "cond": sp.lambdify(args, cond / rT, "numpy"),
        "dA": sp.lambdify(args, [dA.dot(S.j), dA.dot(S.k), dA.dot(S.i)], "numpy"),
        "ds_ohm": sp.lambdify(args, ds_ohm, "numpy"),
        "dev": {k: sp.lambdify(args, v, "numpy") for k, v in dev.items()},
    }
    return out


def evaluate(model, fields=default_fields):
    """(state, expected Laplacian contribution (8 fields) with the reference's departures,
    the departures themselves (8 fields each), the thermal-conduction part of ds)."""
    fn = build(model, fields)
    Rg, Pg, Tg = np.meshgrid(model.r, model.phi, model.theta, indexing="ij")
    shp = Rg.shape
    bc = lambda v: np.broadcast_to(np.asarray(v, dtype=np.float64), shp)
    state = np.stack([bc(v) for v in fn["fields"](Rg, Tg, Pg)])
    exp = np.zeros_like(state)
    f = [bc(v) for v in fn["force"](Rg, Tg, Pg)]
    exp[1], exp[2], exp[3] = f
    exp[4] = bc(fn["ds"](Rg, Tg, Pg)) + bc(fn["ds_ohm"](Rg, Tg, Pg))
    exp[5:8] = np.stack([bc(v) for v in fn["dA"](Rg, Tg, Pg)])
    devs = {}
    for k, g in fn["dev"].items():
        d = np.zeros_like(state)
        d[3 if k == "du" else 4] = bc(g(Rg, Tg, Pg))
        devs[k] = d
    cond = bc(fn["cond"](Rg, Tg, Pg))
    return state, exp, devs, cond
