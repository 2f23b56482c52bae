"""Pins the wall ghost rules of the oracle's SLD diffusion (Diffusion_R diffusion.f90:582-727,
Diffusion_Th :1008-1120).

The wall faces themselves carry no flux (``delue(:,1) = 0``, ``delue(:,mynr+1) = 0``,
diffusion.f90:805-831), so a wall's ghosts act only through the ghost g next to the wall
(``b1(:,4)`` / ``b2(:,1)``), in the Van Albada slope of the first interior face.  For a field
that is linear in the wall-normal index, t_k = a + q k counted from the wall, every other face
is flux free and that face carries

    delue = (q - minmod(a - g, q)) / 2            (Van Albada with |q| >> tinyvar is minmod)

times the speed and geometry of that face.  The ratio of the wall cell's tendency at two offsets
a is therefore a closed form in the ghost rule alone:

    D(a) / D(a_ref) = (q - minmod(a - g(a), q)) / (q - minmod(a_ref - g(a_ref), q)).

The rules, by variable (r walls; theta walls):
    rho   zero; copy                 w  extrapolated in t/r; zero
    v     extrapolated in t/r; extrapolated in t/sin(theta)
    u     copy; copy                 s  zero (constant flux under explicit diffusion: a
                                        linear ghost from the luminosity); copy
    A_th  copy (zero when perfectly conducting); zero under Theta_Normal_Field, else copy
    A_ph  copy (zero when perfectly conducting); zero under Theta_Perf_Conductor, else copy
    A_r   copy; zero under Theta_Perf_Conductor, else copy
(rho diffuses only under Mass_Diffusion, whose corrections change the other variables only)
where "copy" is g = the wall value (no flux: the ratio is 1 for every a), "zero" is g = 0
(ratio (q - a)/(q - a_ref)) and "extrapolated" is g = -10/3 t_0/x_0 + 6 t_1/x_1 - 2 t_2/x_2
+ 1/3 t_3/x_3 with x = r or sin(theta) of the first four points.  Swapping any two rules, or
the x of an extrapolation, changes the ratio by far more than the tolerance.  The cell next to
the wall cell takes the opposite flux: V_0 D_0 = -V_1 D_1.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from cssx.model import Config, Model

TINY = [1e-8, 1e-8, 1e-8, 1e-8, 1e-3, 1.0, 1.0, 1.0]    # tinyvar (diffusion.f90:8-11, :95-104)
NAMES = ["rho", "w", "v", "u", "s", "A_theta", "A_phi", "A_r"]


def _minmod(a, b):
    return np.sign(b) * min(abs(a), abs(b)) if a * b > 0 else 0.0


def _model(**kw):
    base = dict(nr=16, nth=16, nphi=8, sld_coef=0.4, dstype=0, Mass_Diffusion=False,
                Do_Viscous_Heating=False, Do_Ohmic_Heating=False)
    base.update(kw)
    return Model(Config(**base))


def _state(m, iv, axis, side, a, q):
    """Background state plus, in variable iv, the field a + q k with k counted from the chosen
    wall along the axis, constant in the other two directions."""
    c = m.c
    st = np.zeros((c.nv, c.nr, c.nphi, c.nth))
    st[0] = m.rho_bar[:, None, None]
    st[4] = m.s_bar[:, None, None]
    n = c.nr if axis == "r" else c.nth
    k = np.arange(n) if side == "lo" else np.arange(n)[::-1]
    f = a + q * k
    st[iv] += f[:, None, None] if axis == "r" else f[None, None, :]
    return st


def _cells(m, axis, side):
    """(wall cell, next cell) indices into D[iv] away from the other walls' rows."""
    c = m.c
    if axis == "r":
        i0, i1 = (0, 1) if side == "lo" else (c.nr - 1, c.nr - 2)
        return (i0, 0, c.nth // 2), (i1, 0, c.nth // 2)
    i0, i1 = (0, 1) if side == "lo" else (c.nth - 1, c.nth - 2)
    return (c.nr // 2, 0, i0), (c.nr // 2, 0, i1)


def _ghost(m, rule, axis, side, a, q):
    """g next to the wall for the field a + q k (k from the wall)."""
    c = m.c
    t = a + q * np.arange(4)
    if rule == "copy":
        return a
    if rule == "zero":
        return 0.0
    if rule == "extrap":
        if axis == "r":
            x = m.r[:4] if side == "lo" else m.r[::-1][:4]
        else:
            x = np.sin(m.theta[:4]) if side == "lo" else np.sin(m.theta[::-1][:4])
        return -10.0 / 3.0 * t[0] / x[0] + 6.0 * t[1] / x[1] - 2.0 * t[2] / x[2] + t[3] / (3.0 * x[3])
    if rule == "flux":
        # diffusion.f90:614-618 (bottom: + dr_ind(1) lumr1 / (4 pi r1^2 ks(1) dxdr(1))) and
        # :688-692 (top: - dr_ind(1) lumr2 / (4 pi r2^2 ks(nr) dxdr(nr)))
        if side == "lo":
            return a + m.dr[0] * c.lumr1 / (4 * np.pi * m.r[0] ** 2 * m.ks[0] * m.dxdr[0])
        return a - m.dr[0] * c.lumr2 / (4 * np.pi * m.r[-1] ** 2 * m.ks[-1] * m.dxdr[-1])
    raise ValueError(rule)


def _response(m, rule, axis, side, a, q):
    return q - _minmod(a - _ghost(m, rule, axis, side, a, q), q)


def _probes(m, rule, axis, side, Q):
    """(a, q) pairs that put a - g(a) of one rule at chosen points of the limiter's range (the
    test runs every rule's probes, so each alternative rule is probed where it responds).  a - g is linear
    in a for a fixed q (L(a) = alpha a + gamma); where alpha = 0 (copy, constant flux) the slope
    q is varied instead."""
    L = lambda a, q: a - _ghost(m, rule, axis, side, a, q)   # noqa: E731
    alpha = L(1.0, Q) - L(0.0, Q)
    out = []
    if abs(alpha) > 1e-12:
        gamma = L(0.0, Q)
        for frac in (0.2, 0.5, 0.8, 1.5, -0.3):
            out.append(((frac * Q - gamma) / alpha, Q))
    else:
        delta = -L(0.0, Q)            # g - a, the same for every a (0 for copy)
        if delta == 0.0:
            out = [(0.3 * Q, Q), (-0.7 * Q, 2 * Q), (2.0 * Q, -Q), (0.1 * Q, 0.5 * Q)]
        else:
            out = [(0.3 * Q, -delta / f) for f in (0.2, 0.5, 0.8)] + [(0.3 * Q, Q), (-Q, delta / 0.5)]
    return out


def _volumes(m):
    c = m.c
    r, dr, th = m.r, m.dr, m.theta
    rl = np.concatenate([[r[0] - 0.5 * dr[0]], 0.5 * (r[1:] + r[:-1])])
    rh = np.concatenate([0.5 * (r[1:] + r[:-1]), [r[-1] + 0.5 * dr[-1]]])
    dth = (th[-1] - th[0]) / (c.nth - 1)
    tl = np.concatenate([[th[0] - 0.5 * dth], 0.5 * (th[1:] + th[:-1])])
    thh = np.concatenate([0.5 * (th[1:] + th[:-1]), [th[-1] + 0.5 * dth]])
    return ((rh ** 3 - rl ** 3) / 3.0)[:, None, None] * (np.cos(tl) - np.cos(thh))[None, None, :] * m.dphi


# (variable, axis, rule, model flags)
CASES = [
    (0, "r", "zero", {"Mass_Diffusion": True}), (1, "r", "extrap", {}), (2, "r", "extrap", {}), (3, "r", "copy", {}),
    (4, "r", "zero", {}), (5, "r", "copy", {}), (6, "r", "copy", {}), (7, "r", "copy", {}),
    (5, "r", "zero", {"Bottom_Perfect_Cond": True, "Top_Perfect_Cond": True}),
    (6, "r", "zero", {"Bottom_Perfect_Cond": True, "Top_Perfect_Cond": True}),
    (4, "r", "flux", {"SLaplacian": True, "Bottom_Constant_Flux": True, "Bottom_Constant_Temp": False,
                      "Top_Constant_Flux": True, "Top_Constant_Temp": False, "lumr1": 30.0, "lumr2": 20.0}),
    (0, "theta", "copy", {"Mass_Diffusion": True}), (1, "theta", "zero", {}), (2, "theta", "extrap", {}), (3, "theta", "copy", {}),
    (4, "theta", "copy", {}),
    (5, "theta", "zero", {"Theta_Normal_Field": True}), (5, "theta", "copy", {"Theta_Normal_Field": False}),
    (6, "theta", "zero", {"Theta_Normal_Field": False, "Theta_Perf_Conductor": True}),
    (6, "theta", "copy", {"Theta_Normal_Field": False}),
    (7, "theta", "zero", {"Theta_Normal_Field": False, "Theta_Perf_Conductor": True}),
    (7, "theta", "copy", {"Theta_Normal_Field": False}),
]


@pytest.mark.parametrize("side", ["lo", "hi"])
@pytest.mark.parametrize("iv,axis,rule,flags", CASES,
                         ids=[f"{NAMES[c[0]]}-{c[1]}-{c[2]}{'-' + '-'.join(c[3]) if c[3] else ''}" for c in CASES])
def test_wall_ghost_rule(iv, axis, rule, flags, side):
    m = _model(**flags)
    Q = 1e4 * TINY[iv]          # |q| >> tinyvar, and well above the background's rounding
    c0, c1 = _cells(m, axis, side)
    V = _volumes(m)

    def tendency(a, q):
        D = oracle.diffusion(m, _state(m, iv, axis, side, a, q))
        return D[iv][c0], D[iv][c1]

    rules = ["zero", "copy", "extrap"] + (["flux"] if rule == "flux" else [])
    probes = _probes(m, rule, axis, side, Q)
    probes += [p for o in rules if o != rule for p in _probes(m, o, axis, side, Q)]
    if rule == "flux":
        assert abs(_ghost(m, rule, axis, side, 0.0, Q)) > 100 * TINY[iv]
    a_ref, q_ref = probes[0]
    d_ref, d1_ref = tendency(a_ref, q_ref)
    r_ref = _response(m, rule, axis, side, a_ref, q_ref)
    assert d_ref != 0.0 and r_ref != 0.0
    assert abs(V[c0] * d_ref + V[c1] * d1_ref) <= 1e-10 * abs(V[c0] * d_ref)
    others = [o for o in rules if o != rule]
    moved = {o: 0.0 for o in others}
    for a, q in probes[1:]:
        d, _ = tendency(a, q)
        want = _response(m, rule, axis, side, a, q) / r_ref
        assert abs(d / d_ref - want) < 1e-9 * max(1.0, abs(want)), (NAMES[iv], axis, side, rule, a, q, d / d_ref, want)
        for o in others:
            r0 = _response(m, o, axis, side, a_ref, q_ref)
            if r0 == 0.0:
                moved[o] = np.inf
                continue
            moved[o] = max(moved[o], abs(_response(m, o, axis, side, a, q) / r0 - want))
    # the probes separate this rule from every other one
    for o, mv in moved.items():
        assert mv > 1e-3, (NAMES[iv], axis, side, rule, o, mv)
