"""Manufactured solution for the MHD right-hand side.

Smooth analytic fields (rho, u, s, A) on the spherical segment; the continuous right-hand
side is built with sympy.vector's spherical-coordinate operators (gradient, divergence,
curl with Lame coefficients 1, r, r sin(theta)) from the textbook equations the reference
discretizes (rhs_mpi.f90:786-1855, 2342-2346):
  d rho/dt = -div(rho u)
  d u/dt   = -grad(|u|^2/2) + u x curl(u) - grad(P)/rho - g(r) r_hat - 2 Omega x u
             + Omega^2 varpi + curl(B) x B / (4 pi rho),      B = curl A
  d s/dt   = -u . grad(s)
  d A/dt   = u x B
with P = rho^gamma exp(s/Cv) (so grad(P)/rho = pconst T grad(rho)/rho + (gamma-1) T grad(s)).
Every term is kept separately so a test can show each one matters.
"""
from __future__ import annotations

import numpy as np
import sympy as sp
from sympy.vector import CoordSys3D, curl, divergence, gradient


def build(gamma, Cp, omega, GM, amp=1.0):
    S = CoordSys3D("S", transformation="spherical", variable_names=("r", "theta", "phi"))
    r, th, ph = S.r, S.theta, S.phi
    Cv = Cp / gamma
    a = sp.Float(amp)
    rho = 1 + sp.Rational(3, 10) * sp.cos(5 * r) * (2 + sp.sin(2 * th + sp.Rational(3, 10)) * sp.cos(4 * ph))
    s = -sp.Rational(1, 2) + sp.Rational(1, 20) * sp.sin(4 * r) * sp.cos(3 * th) * sp.sin(4 * ph + sp.Rational(1, 5))
    ur = a * sp.Rational(1, 10) * sp.cos(3 * r) * sp.sin(th) * sp.cos(4 * ph)
    ut = a * sp.Rational(1, 20) * sp.sin(2 * r) * sp.cos(2 * th) * sp.sin(4 * ph)
    up = a * sp.Rational(7, 100) * sp.cos(r) * sp.sin(3 * th) * sp.cos(4 * ph + sp.Rational(1, 2))
    Ar = a * sp.Rational(1, 10) * sp.sin(r) * sp.sin(th) * sp.cos(4 * ph + 1)
    At = a * sp.Rational(1, 10) * sp.sin(2 * r) * sp.cos(th) * sp.cos(4 * ph)
    Ap = a * sp.Rational(2, 10) * sp.cos(3 * r) * sp.sin(2 * th) * sp.sin(4 * ph)
    U = ur * S.i + ut * S.j + up * S.k
    A = Ar * S.i + At * S.j + Ap * S.k
    B = curl(A)
    P = rho ** gamma * sp.exp(s / Cv)
    Om = omega * (sp.cos(th) * S.i - sp.sin(th) * S.j)          # Omega z_hat
    varpi = r * sp.sin(th) * (sp.sin(th) * S.i + sp.cos(th) * S.j)
    terms = {
        "advection": -gradient(U.dot(U) / 2) + U.cross(curl(U)),
        "pressure": -gradient(P) / rho,
        "gravity": -(GM / r ** 2) * S.i,
        "coriolis": -2 * Om.cross(U),
        "centrifugal": omega ** 2 * varpi,
        "lorentz": curl(B).cross(B) / (4 * sp.pi * rho),
    }
    comps = {}
    for name, vec in terms.items():
        comps[name] = [vec.dot(S.j), vec.dot(S.k), vec.dot(S.i)]      # (theta, phi, r) = w, v, u slots
    drho = -divergence(rho * U)
    ds = -U.dot(gradient(s))
    dA = U.cross(B)
    fields = [rho, ut, up, ur, s, At, Ap, Ar]
    f_num = sp.lambdify((r, th, ph), fields, "numpy")
    rhs_scalar = sp.lambdify((r, th, ph), [drho, ds, dA.dot(S.j), dA.dot(S.k), dA.dot(S.i)], "numpy")
    term_num = {k: sp.lambdify((r, th, ph), v, "numpy") for k, v in comps.items()}
    b_num = sp.lambdify((r, th, ph), [B.dot(S.j), B.dot(S.k), B.dot(S.i)], "numpy")
    return f_num, rhs_scalar, term_num, b_num


def evaluate(model, amp=1.0):
    """Returns (state, exact_rhs, per-term momentum arrays, exact B) on the model's grid."""
    c = model.c
    f_num, rhs_scalar, term_num, b_num = build(c.gamma, c.Cp, c.omega0, model.GM, amp)
    R, P, T = np.meshgrid(model.r, model.phi, model.theta, indexing="ij")
    shp = R.shape
    bc = lambda v: np.broadcast_to(np.asarray(v, dtype=np.float64), shp)
    state = np.stack([bc(v) for v in f_num(R, T, P)])
    drho, ds, dat, dap, dar = (bc(v) for v in rhs_scalar(R, T, P))
    terms = {k: np.stack([bc(v) for v in fn(R, T, P)]) for k, fn in term_num.items()}
    mom = sum(terms.values())
    exact = np.stack([drho, mom[0], mom[1], mom[2], ds, dat, dap, dar])
    B = np.stack([bc(v) for v in b_num(R, T, P)])
    return state, exact, terms, B
