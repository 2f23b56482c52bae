"""Shell-average, azimuthal-average and shell-slice records in the reference's file formats.

The reference writes each output as one record of a direct-access unformatted file of REAL*8
(``Open(..., form='unformatted', access='direct', recl=...)``; ``Write(unit, rec=k)``), record k
at byte offset (k-1) * (record size), the array in Fortran (column-major) order:

  Shell_Average_Output (report_mpi.f90:783-1094)    averages(nr, nq+2)
      (1,1) nr, (2,1) nq, (3,1) r1, (4,1) r2, (5:nq+4,1) quantity codes; column 2 = r_ind;
      columns 3..nq+2 the theta-phi means of the quantities
  Azimuthal_Average_Output (:441-781)               averages(nth, nr+2, nq)
      (1..7,1,1) nth, nr, nq, r1, r2, th1, th2; (8:nq+7,1,1) codes; (1:nr,2,1) = r_ind;
      (:, 3..nr+2, q) the phi means
  Shell_Slice_Output (:1096-1291)                   averages(nth, nphi+1, nradii, nq)
      (1..11,1,1,1) nradii, nr, nth, nphi, nq, r1, r2, th1, th2, phi1, phi2; then the quantity
      codes and the (1-based) radius indices; (:, 2..nphi+1, k, q) the slice values

CSS-Python reads them back with read_shell_avg.py, read_az_avg.py and read_shell_slice.py.
The quantities are formed on the GPU from an FP64 state (A in slots 6-8) and B = curl A with
the Compute_B wall conditions, as the reference's report layer sees vars0 after Compute_B(0).
The Do_SLD_Output quantities (SLD fluxes at cell centres: shell 19, azimuthal 18, 19, 24, 25)
are not formed here (ValueError).
"""
from __future__ import annotations

import math
import os

import numpy as np
import torch

from .mhd import Solver

SHELL_QUANTITIES = tuple(range(1, 19))
AZ_QUANTITIES = tuple(range(1, 18)) + (20, 21, 22, 23)
SLICE_QUANTITIES = (1, 2, 3, 4, 5, 6, 7, 8, 10, 11, 12, 13, 14)


class _Fields:
    """The point quantities of the report layer for one FP64 state, on the GPU."""

    def __init__(self, sol: Solver, y: torch.Tensor):
        m = sol.model
        self.sol, self.m, self.c = sol, m, m.c
        self.rho, self.w, self.v, self.u, self.s = y[0], y[1], y[2], y[3], y[4]
        # temperature as Compute_Derivatives forms it (report_mpi.f90:408)
        self.T = self.rho.pow(m.gamm1) * torch.exp(self.s / m.Cv) * m.invtconst
        self.P = (m.gamm1 * m.Cv) * self.rho * self.T
        if self.c.magnetic:
            _, B = sol.rhs(y, want_b=True)
            self.bt, self.bp, self.br = B[0], B[1], B[2]
        else:
            self.bt = self.bp = self.br = None
        self._y = y
        self._dsdr = None

    def usqr(self):
        return 0.5 * self.rho * (self.u * self.u + self.v * self.v + self.w * self.w)

    def bsqr(self):
        self._need_b()
        return (self.br * self.br + self.bt * self.bt + self.bp * self.bp) * (0.125 / math.pi)

    def bpol(self):
        self._need_b()
        return torch.sqrt(self.br * self.br + self.bt * self.bt)

    def _need_b(self):
        if self.br is None:
            raise ValueError("a magnetic quantity of a hydrodynamic run")

    def dsdr(self):
        """ds/dr by the compact operator with the entropy wall conditions of Compute_Derivatives
        (report_mpi.f90:290-320): a constant-flux wall takes dsdr1/dsdr2 / dxdr."""
        if self._dsdr is None:
            from .operators import Operators
            m, c = self.m, self.c
            ops = Operators(c.nr, c.nth, c.nphi, m.dri, m.dti, m.dpi)
            dev = self.s.device
            glo = torch.zeros((4, c.nphi, c.nth), dtype=torch.float64, device=dev)
            ghi = torch.zeros_like(glo)
            lbcr_s = m.ibcr[5]
            if lbcr_s in (2, 4):
                glo[0] = torch.as_tensor(m.dsdr1 / m.dxdr[0], device=dev)[None, :]
            if lbcr_s in (3, 4):
                ghi[0] = torch.as_tensor(m.dsdr2 / m.dxdr[-1], device=dev)[None, :]
            d = ops.dbydr(self.s.contiguous(), lbcr_s, glo, ghi)
            ops.close()
            self._dsdr = d * torch.as_tensor(m.dxdr, device=dev)[:, None, None]
        return self._dsdr


def _shell_mean(x):
    """theta-phi mean (Sum / (nth nphi)) of an (nr, nphi, nth) field -> (nr,)"""
    return x.mean(dim=(1, 2))


def shell_means(F: _Fields) -> dict:
    """avgT, avgrho, avgP, avgvr, avgS of Shell_Average_Output (report_mpi.f90:815-840)."""
    return {"T": _shell_mean(F.T), "rho": _shell_mean(F.rho), "P": _shell_mean(F.P), "vr": _shell_mean(F.u),
            "S": _shell_mean(F.s)}


def shell_average_values(sol: Solver, y: torch.Tensor, quantities, k0=None, _fields=None) -> np.ndarray:
    """(nr, nq) theta-phi means of the Shell_Average_Output quantities (report_mpi.f90:880-985).
    k0: the radiative flux profile k0(r) of quantity 15 (zeros when not given)."""
    F = _fields or _Fields(sol, y)
    c, m = F.c, F.m
    av = shell_means(F)
    cols = []
    for q in quantities:
        q = int(q)
        if q == 1:
            x = av["rho"]
        elif q == 2:
            x = av["T"]
        elif q == 3:
            x = av["S"]
        elif q == 4:
            x = av["P"]
        elif q == 5:
            x = av["vr"]
        elif q == 6:
            x = _shell_mean(F.w)
        elif q == 7:
            x = _shell_mean(F.v)
        elif q == 8:
            x = _shell_mean(F.usqr())
        elif q == 9:
            x = _shell_mean(F.dsdr())
        elif q in (10, 11, 12):
            F._need_b()
            x = _shell_mean({10: F.br, 11: F.bt, 12: F.bp}[q])
        elif q == 13:
            x = _shell_mean(F.bsqr())
        elif q == 14:
            x = _shell_mean(F.bpol())
        elif q == 15:
            # non-radiative branch (:948-953), one subdomain (nnp nnt = 1)
            kk = torch.zeros(c.nr, dtype=torch.float64, device=F.s.device) if k0 is None else \
                torch.as_tensor(np.asarray(k0, np.float64), device=F.s.device)
            x = kk
            if c.Laplacian or c.SLaplacian:
                x = -torch.as_tensor(m.ks, device=F.s.device) * _shell_mean(F.dsdr()) + kk
        elif q == 16:
            x = c.Cp * _shell_mean((F.u - av["vr"][:, None, None]) * F.rho * (F.T - av["T"][:, None, None]))
        elif q == 17:
            x = _shell_mean(F.usqr() * (F.u - av["vr"][:, None, None]))
        elif q == 18:
            x = _shell_mean((F.P - av["P"][:, None, None]) * (F.u - av["vr"][:, None, None]))
        else:
            raise ValueError(f"shell-average quantity {q} is not formed here")
        cols.append(x)
    return torch.stack(cols, dim=1).cpu().numpy()


def shell_average_record(sol: Solver, y: torch.Tensor, quantities, k0=None) -> np.ndarray:
    """averages(nr, nq+2) of Shell_Average_Output, header included (:1005-1011)."""
    c, m = sol.model.c, sol.model
    nq = len(quantities)
    if c.nr < nq + 4:
        raise ValueError("the header column needs nr >= nq + 4")
    rec = np.zeros((c.nr, nq + 2))
    rec[:, 2:] = shell_average_values(sol, y, quantities, k0)
    rec[0, 0], rec[1, 0], rec[2, 0], rec[3, 0] = c.nr, nq, c.r1, c.r2
    rec[4:nq + 4, 0] = quantities
    rec[:, 1] = m.r
    return rec


def _phi_mean(x):
    """Sum over phi / nphi of an (nr, nphi, nth) field -> (nth, nr)"""
    return x.mean(dim=1).t()


def azimuthal_average_values(sol: Solver, y: torch.Tensor, quantities, _fields=None) -> np.ndarray:
    """(nth, nr, nq) phi means of the Azimuthal_Average_Output quantities (:478-733).  The
    fluctuations of quantities 1-4 are taken about the shell means of the same state, which
    Shell_Average_Output leaves in avgrho, avgT, avgS, avgP."""
    F = _fields or _Fields(sol, y)
    c, m = F.c, F.m
    dev = F.s.device
    av = shell_means(F)
    nphi = c.nphi
    vs = {"rho": F.rho.mean(dim=1), "w": F.w.mean(dim=1), "v": F.v.mean(dim=1), "u": F.u.mean(dim=1)}
    if c.magnetic:
        vs.update({"bt": F.bt.mean(dim=1), "bp": F.bp.mean(dim=1), "br": F.br.mean(dim=1)})
    r = torch.as_tensor(m.r, device=dev)[:, None]
    sn = torch.sin(torch.as_tensor(m.theta, device=dev))[None, :]
    rs = r * sn                                      # (nr, nth)
    cols = []
    for q in quantities:
        q = int(q)
        if q == 1:
            x = (F.rho - av["rho"][:, None, None]).mean(dim=1)
        elif q == 2:
            x = (F.T - av["T"][:, None, None]).mean(dim=1)
        elif q == 3:
            x = (F.s - av["S"][:, None, None]).mean(dim=1)
        elif q == 4:
            x = (F.P - av["P"][:, None, None]).mean(dim=1)
        elif q in (5, 6, 7):
            x = vs[{5: "u", 6: "w", 7: "v"}[q]]
        elif q == 8:
            x = F.usqr().mean(dim=1)
        elif q in (9, 10, 11):
            F._need_b()
            x = vs[{9: "br", 10: "bt", 11: "bp"}[q]]
        elif q == 12:
            x = F.bsqr().mean(dim=1)
        elif q == 13:
            x = F.bpol().mean(dim=1)
        elif q in (14, 15):
            vel = F.u if q == 14 else F.w
            x = rs * (F.rho * vel).sum(dim=1) * (vs["v"] + rs * c.omega0) / nphi
        elif q in (16, 17):
            vel, vbar = (F.u, vs["u"]) if q == 16 else (F.w, vs["w"])
            x = rs * vs["rho"] * ((vel - vbar[:, None, :]) * (F.v - vs["v"][:, None, :])).sum(dim=1) / nphi
        elif q in (20, 21):
            F._need_b()
            x = rs / (4 * math.pi) * vs["br" if q == 20 else "bt"] * vs["bp"]
        elif q in (22, 23):
            # the reference sums over phi without dividing by nphi here (:703-717)
            F._need_b()
            b, bbar = (F.br, vs["br"]) if q == 22 else (F.bt, vs["bt"])
            x = rs / (4 * math.pi) * ((b - bbar[:, None, :]) * (F.bp - vs["bp"][:, None, :])).sum(dim=1)
        else:
            raise ValueError(f"azimuthal-average quantity {q} is not formed here")
        cols.append(x.t())                           # (nth, nr)
    return torch.stack(cols, dim=2).cpu().numpy()


def azimuthal_average_record(sol: Solver, y: torch.Tensor, quantities) -> np.ndarray:
    """averages(nth, nr+2, nq) of Azimuthal_Average_Output, header included (:750-759)."""
    c, m = sol.model.c, sol.model
    nq = len(quantities)
    if c.nth < max(nq + 7, c.nr):
        raise ValueError("the header row needs nth >= nq + 7 and nth >= nr (r_ind is stored along theta)")
    rec = np.zeros((c.nth, c.nr + 2, nq))
    rec[:, 2:, :] = azimuthal_average_values(sol, y, quantities)
    rec[0:7, 0, 0] = [c.nth, c.nr, nq, c.r1, c.r2, c.th1, c.th2]
    rec[7:nq + 7, 0, 0] = quantities
    rec[0:c.nr, 1, 0] = m.r
    return rec


def shell_slice_record(sol: Solver, y: torch.Tensor, quantities, radii) -> np.ndarray:
    """averages(nth, nphi+1, nradii, nq) of Shell_Slice_Output (:1130-1219); radii are 1-based
    radius indices."""
    c = sol.model.c
    nq, nrad = len(quantities), len(radii)
    if c.nth < 11 + nq + nrad:
        raise ValueError("the header row needs nth >= 11 + nq + nradii")
    if any(int(k) < 1 or int(k) > c.nr for k in radii):
        raise ValueError("radius indices are 1-based, 1..nr")
    F = _Fields(sol, y)
    idx = torch.as_tensor([int(k) - 1 for k in radii], device=F.s.device)
    rec = np.zeros((c.nth, c.nphi + 1, nrad, nq))
    for j, q in enumerate(quantities):
        q = int(q)
        field = {1: lambda: F.rho, 2: lambda: F.T, 3: lambda: F.s, 4: lambda: F.P, 5: lambda: F.u, 6: lambda: F.w,
                 7: lambda: F.v, 8: F.usqr, 10: lambda: F.br, 11: lambda: F.bt, 12: lambda: F.bp, 13: F.bsqr,
                 14: F.bpol}.get(q)
        if field is None:
            raise ValueError(f"shell-slice quantity {q} is not formed here")
        x = field()
        if x is None:
            raise ValueError("a magnetic quantity of a hydrodynamic run")
        sl = x.index_select(0, idx)                  # (nradii, nphi, nth)
        rec[:, 1:, :, j] = sl.permute(2, 1, 0).cpu().numpy()
    rec[0:11, 0, 0, 0] = [nrad, c.nr, c.nth, c.nphi, nq, c.r1, c.r2, c.th1, c.th2, c.phi1, c.phi2]
    rec[11:11 + nq, 0, 0, 0] = quantities
    rec[11 + nq:11 + nq + nrad, 0, 0, 0] = radii
    return rec


def write_record(path: str, rec: np.ndarray, record: int) -> None:
    """Write rec as record `record` (1-based) of a direct-access REAL*8 file, as the reference's
    ``Write(unit, rec=record) averages`` does: byte offset (record-1) * rec.nbytes, Fortran order."""
    if record < 1:
        raise ValueError("records are numbered from 1")
    data = np.asarray(rec, dtype="<f8").tobytes(order="F")
    mode = "r+b" if os.path.exists(path) else "w+b"
    with open(path, mode) as f:
        f.seek((record - 1) * len(data))
        f.write(data)
