"""Run diagnostics of the reference's report layer, computed on the GPU.

  scalar_quantities   Scalar_Output (report_mpi.f90:173-262): kinetic energy and its radial,
                      meridional and differential-rotation parts, the largest Mach number, the
                      magnetic energy and the largest |B|, as the reference sums them (plain
                      point sums over the grid divided by the point count)
  scalar_line         one line of scalar.dat in the reference's (i10,9d15.7) / (i10,7d15.7) format
  append_scalar       appends that line to a file
  conservation        the totals Compute_Step_Size prints (rhs_mpi.f90:156-216): E, L, m, m r sin
  azimuthal_average   the phi mean of fields, the core of Azimuthal_Average_Output
                      (report_mpi.f90:441-781)
Shell averages with the reference's Simpson weights are cssx.analysis.shell_avg.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .mhd import Solver


def scalar_quantities(sol: Solver, y: torch.Tensor) -> dict:
    """Scalar_Output of an FP64 state (A in slots 6-8; B = curl A as Compute_B(0) leaves it)."""
    m = sol.model
    c = m.c
    N = c.npts
    rho, w, v, u, s = y[0], y[1], y[2], y[3], y[4]
    # the "temperature" Scalar_Output forms: (gamma-1) Cp T, a squared sound speed
    tsq = rho.pow(m.gamm1) * torch.exp(s / m.Cv) * (m.invtconst * m.gamm1 * c.Cp)
    usq = w * w + v * v + u * u
    q = {
        "kinener": (rho * usq).sum().item() / (2 * N),
        "cke": (rho * u * u).sum().item() / (2 * N),
        "drke": (rho * v * v).sum().item() / (2 * N),
    }
    q["mcke"] = q["cke"] + (rho * w * w).sum().item() / (2 * N)
    if c.magnetic:
        _, B = sol.rhs(y, want_b=True)
        bsq = (B * B).sum(0)
        q["magener"] = bsq.sum().item() / (8 * math.pi * N)
        q["bmax"] = bsq.max().sqrt().item() * 2.0 * math.sqrt(2.0 * math.pi)
        q["umach"] = math.sqrt((usq / (2.0 * bsq / rho + usq + tsq)).max().item())
    else:
        q["umach"] = math.sqrt((usq / (usq + tsq)).max().item())
    return q


def fortran_d(x: float, w: int = 15, d: int = 7) -> str:
    """Fortran Dw.d edit descriptor (0.ddddddd with a D exponent; three-digit exponents drop
    the letter), right-justified."""
    x = float(x)
    if not math.isfinite(x):
        return ("NaN" if math.isnan(x) else ("-Infinity" if x < 0 else "Infinity")).rjust(w)
    sign = "-" if (x < 0 or (x == 0 and math.copysign(1.0, x) < 0)) else ""
    a = abs(x)
    if a == 0.0:
        mant, e = "0" * d, 0
    else:
        digits, ex = f"{a:.{d - 1}e}".split("e")
        mant = digits.replace(".", "")
        e = int(ex) + 1
    es = f"D{'+' if e >= 0 else '-'}{abs(e):02d}" if abs(e) <= 99 else f"{'+' if e >= 0 else '-'}{abs(e):03d}"
    body = f"{sign}0.{mant}{es}"
    if len(body) > w:
        body = f"{sign}.{mant}{es}"
    return body.rjust(w) if len(body) <= w else "*" * w


def scalar_line(istep: int, sim_time: float, deltat: float, q: dict, magnetic: bool = True) -> str:
    """scalar.dat line (report_mpi.f90:253-257)."""
    keys = ["kinener", "cke", "mcke", "drke", "umach"] + (["magener", "bmax"] if magnetic else [])
    vals = [sim_time, deltat] + [q[k] for k in keys]
    s = str(int(istep)).rjust(10)
    return s + "".join(fortran_d(v) for v in vals)


def append_scalar(path: str, istep: int, sim_time: float, deltat: float, q: dict, magnetic: bool = True) -> None:
    with open(path, "a") as f:
        f.write(scalar_line(istep, sim_time, deltat, q, magnetic) + "\n")


def conservation(sol: Solver, y: torch.Tensor) -> dict:
    return sol.conservation(y)


def azimuthal_average(fields: torch.Tensor) -> np.ndarray:
    """phi mean of (nq, nr, nphi, nth) fields -> (nq, nr, nth)."""
    return fields.double().mean(dim=2).cpu().numpy()
