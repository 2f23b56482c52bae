"""The plain-derivative workload's field: one seeded scalar on a 64 x 64 x 128 spherical segment.

r in [0.72, 0.97] on the tanh-stretched map r(x) = (r1+r2)/2 + (r2-r1)/2 tanh(k(2x-1))/tanh(k),
k = 1.5, x uniform on [0, 1]; theta uniform on [pi/4, 3pi/4] (both ends on the grid); phi
uniform on [0, pi/2) with spacing (pi/2)/nphi.  f = sum over 8 terms of
a_k sin(kr_k r + p_k) sin(mt_k theta) cos(mp_k phi), a_k ~ N(0, 1), integer modes in 1..6,
phases U[0, 2 pi), drawn from numpy's default_rng(1) in that order.

The field is evaluated on four extra phi rows at each end: they are the phi halo buffers a
decomposed rank would receive (layout (nr, 4, nth), row 3 of the lower buffer and row 0 of the
upper one next to the subdomain), since the modes are not periodic on [0, pi/2).
"""
from __future__ import annotations

import numpy as np

NR, NTH, NPHI = 64, 64, 128
R1, R2, STRETCH = 0.72, 0.97, 1.5


def build(seed: int = 1):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(8)
    kr = rng.integers(1, 7, 8)
    mt = rng.integers(1, 7, 8)
    mp = rng.integers(1, 7, 8)
    ph0 = rng.uniform(0.0, 2.0 * np.pi, 8)
    x = np.linspace(0.0, 1.0, NR)
    tk = np.tanh(STRETCH)
    tz = np.tanh(STRETCH * (2.0 * x - 1.0))
    r = 0.5 * (R1 + R2) + 0.5 * (R2 - R1) * tz / tk
    drdx = 0.5 * (R2 - R1) * 2.0 * STRETCH * (1.0 - tz ** 2) / tk
    th = np.linspace(np.pi / 4, 3 * np.pi / 4, NTH)
    dph = (np.pi / 2) / NPHI
    ph = np.arange(-4, NPHI + 4) * dph
    R, P, T = r[:, None, None], ph[None, :, None], th[None, None, :]
    f = np.zeros((NR, NPHI + 8, NTH))
    fr = np.zeros_like(f)
    ft = np.zeros_like(f)
    fp = np.zeros_like(f)
    for i in range(8):
        sr, cr = np.sin(kr[i] * R + ph0[i]), np.cos(kr[i] * R + ph0[i])
        st, ct = np.sin(mt[i] * T), np.cos(mt[i] * T)
        sp, cp = np.sin(mp[i] * P), np.cos(mp[i] * P)
        f += a[i] * sr * st * cp
        fr += a[i] * kr[i] * cr * st * cp
        ft += a[i] * mt[i] * sr * ct * cp
        fp -= a[i] * mp[i] * sr * st * sp
    core = slice(4, NPHI + 4)
    return {
        "f": np.ascontiguousarray(f[:, core]),
        "ghost_lo": np.ascontiguousarray(f[:, :4]),
        "ghost_hi": np.ascontiguousarray(f[:, NPHI + 4:]),
        "fr": np.ascontiguousarray(fr[:, core]),
        "ft": np.ascontiguousarray(ft[:, core]),
        "fp": np.ascontiguousarray(fp[:, core]),
        "dxdr": 1.0 / drdx,
        "dri": float(NR - 1),                 # x spacing 1/(nr-1)
        "dti": (NTH - 1) / (np.pi / 2),
        "dpi": 1.0 / dph,
        "r": r,
    }
