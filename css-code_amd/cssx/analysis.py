"""GPU-backed CSS-Python analysis routines that call compact_fd6.

CSS-Python's analysis scripts differentiate checkpoint data one pencil at a time with the
numpy compact_fd6 (vectorcurl.py:55-153 loops over every (phi, r), (theta, r) and
(theta, phi) pencil).  The functions here keep those scripts' signatures and arithmetic, but
send whole fields through cssx.derivatives.compact_fd6, which solves every pencil of an axis
in one library call on the GPU.

  read_checkpoint   CSS-Python/read_checkpoint.py:20-108 (same return tuple and grids)
  vectorcurl        CSS-Python/vectorcurl.py:20-161 (same component labels: the script calls
                    slot 0 of its data "w --> u_phi" and slot 1 "v --> u_theta", the reverse of
                    the Fortran state's w = u_theta, v = u_phi; the arithmetic follows the
                    script exactly, so the results equal the script's)
  shell_avg         shell_avg.py:13-75 with integrate.py's Simpson rule, as one weighted
                    reduction on the GPU
  rossby_from_chk   rossby_from_checkpoint.py:15-220 ('adv', 'scale', 'ens', 'turn'; no plots)
  slice_vorticity   the radial vorticity of plot_shell_slice.py:449-512 (Vorticity) on shell-slice
                    data, every pencil of a slice record in two library calls
"""
from __future__ import annotations

import os

import numpy as np
import torch

from . import checkpoint as ck
from .derivatives import compact_fd6


def _dev():
    if not torch.cuda.is_available():
        raise RuntimeError("cssx.analysis needs a ROCm GPU")
    return torch.device("cuda")


def read_checkpoint(iter, case, mag_curl=False, vel_curl=False, base_dir=""):
    """(data (nth, nph, nr, nq), rad, theta, phi, time[dt, t], header file, ierr) of
    <base_dir><case>/Checkpoints/<iter:07d>/, as read_checkpoint.py returns them.  With
    mag_curl only bt, bp, br are read; with vel_curl only w, v, u (nq = 3)."""
    d = f"{base_dir}{case}/Checkpoints/{int(iter):07d}/"
    fname = d + "header"
    try:
        h = ck.read_header(fname)
    except (OSError, ValueError):
        return None, None, None, None, None, None, True
    if mag_curl:
        names = ["bt", "bp", "br"]
    elif vel_curl:
        names = ["w", "v", "u"]
    else:
        names = list(ck.NAMES[:h.nv])
    data = np.empty((h.nth, h.nphi, h.nr, len(names)))
    for iv, nm in enumerate(names):
        fn = os.path.join(d, "checkpoint." + nm)
        try:
            a = np.fromfile(fn, dtype="=f8")
        except OSError:
            return None, None, None, None, None, None, True
        if a.size != h.nth * h.nphi * h.nr:
            return None, None, None, None, None, None, True
        data[:, :, :, iv] = a.reshape((h.nth, h.nphi, h.nr), order="F")
    nr, nth, nph = h.nr, h.nth, h.nphi
    # grids as read_checkpoint.py:96-99 builds them (phi with the nph-1 divisor it uses)
    rad = (h.r2 - h.r1) * np.arange(nr) / float(nr - 1) + h.r1
    theta = (h.th2 - h.th1) * np.arange(nth) / float(nth - 1) + h.th1
    phi = (h.phi2 - h.phi1) * np.arange(nph) / float(nph - 1) + h.phi1
    return data, rad, theta, phi, np.array([h.deltat, h.sim_time]), fname, False


def vectorcurl(iter, case, mag=False, test=None, base_dir=""):
    """curl (nth, nphi, nr, 3), rad, theta, phi of checkpoint velocity (or B) data, or of
    test = [data, rad, theta, phi]; vectorcurl.py:20-161 with every pencil loop batched."""
    iphi, ith, irad = 0, 1, 2
    if test is None:
        data, rad, theta, phi, _, header, ierr = read_checkpoint(iter, case, mag_curl=mag, vel_curl=not mag,
                                                                 base_dir=base_dir)
        if ierr:
            raise IOError(f"could not extract data from checkpoint (header file {header})")
    else:
        data, rad, theta, phi = test[0], test[1], test[2], test[3]
    dev = _dev()
    rad = np.asarray(rad, dtype=np.float64)
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    D = torch.as_tensor(np.asarray(data, dtype=np.float64), device=dev)
    nth, nphi, nr = len(theta), len(phi), len(rad)
    dri = 1.0 / (rad[-1] - rad[0])
    dti = 1.0 / (theta[-1] - theta[0])
    dpi = 1.0 / (phi[-1] - phi[0])
    sn = torch.as_tensor(np.sin(theta), device=dev)[:, None, None]
    csc = 1.0 / sn
    rinv = torch.as_tensor(1.0 / rad, device=dev)[None, None, :]
    rr = torch.as_tensor(rad, device=dev)[None, None, :]
    curl = torch.empty((nth, nphi, nr, 3), dtype=torch.float64, device=dev)
    # theta: (ibc 4 with the script's Neumann arguments, which dtype 0 ignores) and ibc 0
    curl[..., irad] = rinv * csc * compact_fd6(dti, (sn * D[..., iphi]).contiguous(), 0.0, 0.0, 4, axis=0)
    curl[..., iphi] = -rinv * compact_fd6(dti, D[..., irad].contiguous(), 0.0, 0.0, 0, axis=0)
    # phi: internal pencils (dtype 3) closed with the periodic neighbours as ghosts
    for comp, sign, dst in ((irad, 1.0, ith), (ith, -1.0, irad)):
        f = D[..., comp].contiguous()
        b1 = f[:, nphi - 4:nphi, :].contiguous()
        b2 = f[:, 0:4, :].contiguous()
        term = sign * rinv * csc * compact_fd6(dpi, f, b1, b2, 4, dtype=3, axis=1)
        if dst == ith:
            curl[..., ith] = term
        else:
            curl[..., irad] += term
    # r
    curl[..., ith] += -rinv * compact_fd6(dri, (rr * D[..., iphi]).contiguous(), 0.0, 0.0, 4, axis=2)
    curl[..., iphi] += rinv * compact_fd6(dri, (rr * D[..., ith]).contiguous(), 0.0, 0.0, 0, axis=2)
    return curl.cpu().numpy(), rad, theta, phi


def simpson_weights(n, dx):
    """Weights w with sum(w*f) = integrate.simpson_integrate(f, dx) (integrate.py:78-103):
    Simpson over pairs of slabs, the last slab of an odd count by dx/12 (-f[n-3] + 8 f[n-2] +
    5 f[n-1])."""
    ns = n - 1
    w = np.zeros(n)
    k = 0
    while k <= ns - 2:
        w[k] += dx / 3.0
        w[k + 1] += 4.0 * dx / 3.0
        w[k + 2] += dx / 3.0
        k += 2
    if ns % 2:
        w[ns - 2] += -dx / 12.0
        w[ns - 1] += 8.0 * dx / 12.0
        w[ns] += 5.0 * dx / 12.0
    return w


def shell_avg(data, theta, phi, numq=None):
    """shell_avg.py:13-75 ('simp'): the sin(theta)-weighted Simpson average over theta and phi of
    data (nth, nphi, nr, nq) (or (nth, nphi, nr) with numq) -> (nr, nq)."""
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    dev = _dev()
    d = torch.as_tensor(np.asarray(data, np.float64), device=dev)
    if numq is not None:
        d = d.reshape(d.shape[0], d.shape[1], d.shape[2], 1)
    area = (phi[-1] - phi[0]) * (np.cos(theta[0]) - np.cos(theta[-1]))
    wt = torch.as_tensor(simpson_weights(len(theta), theta[1] - theta[0]) * np.sin(theta), device=dev)
    wp = torch.as_tensor(simpson_weights(len(phi), phi[1] - phi[0]), device=dev)
    avg = torch.tensordot(wt, torch.tensordot(wp, d, dims=([0], [1])), dims=([0], [0])) / area
    return avg.cpu().numpy()


def rossby_from_chk(iter, case, omega0, method="adv", return_rad=False, base_dir="", gamma=1.525, Cp=3.5e8):
    """Rossby number profile of a checkpoint (rossby_from_checkpoint.py:15-220, without its
    plots): 'adv' U/(2 L Omega0), 'scale' U/(2 H Omega0) with 1/H = -drho/dr / rho, 'ens'
    <|curl v|>/(2 Omega0), 'turn' 2 pi/(tau_c Omega0) with tau_c = 1.4 H_p / U (gamma, Cp as
    the script hard-codes them)."""
    if method == "adv":
        data, radius, theta, phi, _, header, ierr = read_checkpoint(iter, case, vel_curl=True, base_dir=base_dir)
        if ierr:
            raise IOError(f"could not read checkpoint data (header file {header})")
        avg = shell_avg(data, theta, phi)
        U = np.sqrt(avg[:, 2] ** 2)
        ro = U / (2.0 * (radius[-1] - radius[0]) * omega0)
    elif method in ("scale", "turn"):
        data, radius, theta, phi, _, header, ierr = read_checkpoint(iter, case, vel_curl=False, base_dir=base_dir)
        if ierr:
            raise IOError(f"could not read checkpoint data (header file {header})")
        avg = shell_avg(data[..., :5], theta, phi)
        U = np.sqrt(avg[:, 3] ** 2)
        rho = avg[:, 0]
        if method == "scale":
            dri = 1.0 / (radius[-1] - radius[0])
            H = -rho / compact_fd6(dri, rho, 0.0, 0.0, 1, dtype=0)
            ro = U / (2.0 * H * omega0)
        else:
            Cv = Cp / gamma
            dri = 1.0 / (radius[1] - radius[0])     # as the script passes it
            logp = gamma * np.log(rho) + avg[:, 4] / Cv
            H = -1.0 / compact_fd6(dri, logp, 0.0, 0.0, 1, dtype=0)
            ro = 2.0 * np.pi / (1.4 * H / U * omega0)
    elif method == "ens":
        curl, radius, theta, phi = vectorcurl(iter, case, base_dir=base_dir)
        ens = np.sqrt((curl ** 2).sum(axis=3))
        ro = shell_avg(ens, theta, phi, numq=1).reshape(len(radius)) / (2.0 * omega0)
    else:
        raise ValueError(f"unknown method in rossby_from_chk: {method}")
    return (ro, radius) if return_rad else ro


def slice_vorticity(data, theta, phi, r, iut, iup, irec=0):
    """omega_r of shell-slice data (nth, nphi, nradii, nq, nrec) as plot_shell_slice.py:484-510
    forms it: (1/(r sin)) d(sin u_phi)/dtheta - (1/(r sin)) du_theta/dphi, both with the numpy
    compact_fd6 arguments the script passes -- theta: ibc 1, dtype 0; phi: ibc 3, dtype 0, so
    the phi derivative closes with walls (zero slope at the first point), not periodically,
    whatever its comment says; dti, dpi = 1/(max - min) of the grids.  r: the slice radii.
    Returns wr (nth, nphi, nradii)."""
    dev = _dev()
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    r = np.asarray(r, dtype=np.float64)
    D = torch.as_tensor(np.asarray(data, dtype=np.float64)[:, :, :, :, irec], device=dev)
    up = D[:, :, :, iup].contiguous()
    ut = D[:, :, :, iut].contiguous()
    dpi = 1.0 / (phi.max() - phi.min())
    dti = 1.0 / (theta.max() - theta.min())
    sn = torch.as_tensor(np.sin(theta), device=dev)[:, None, None]
    rr = torch.as_tensor(r, device=dev)[None, None, :]
    wr = compact_fd6(dti, (sn * up).contiguous(), 0.0, 0.0, 1, dtype=0, axis=0) / rr / sn
    wr = wr - compact_fd6(dpi, ut, 0.0, 0.0, 3, dtype=0, axis=1) / (rr * sn)
    return wr.cpu().numpy()
