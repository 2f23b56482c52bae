"""Problem setup for one subdomain of the CSS MHD model: grid, background, boundary codes,
initial state.  Host-side FP64 preparation (run once), not part of the step.

Reference behaviour restated here (file:line):
  * uniform theta / periodic phi grids, dphi = (phi2-phi1)/nphi   derivatives.f90:81-84
  * index-space inverse spacings dri/dti/dpi                       derivatives.f90:91-97
  * isentropic polytrope background (rho, T, P, s, g)              mhdsph_mpi.f90:317-361
  * boundary flags -> integer boundary codes per variable          mhdsph_mpi.f90:1302-1462
  * physics constants Cv, pconst, invtconst                        rhs_mpi.f90:59-73
  * entropy window (1-cos^16)(1-cos^16) and A window (1-cos^2)(1-cos^2)
                                                                   mhdsph_mpi.f90:709-715, :1084-1118
The radial grid is the tanh-stretched map of the benchmark configurations (x uniform on [0,1],
r(x) = (r1+r2)/2 + (r2-r1)/2 * tanh(k(2x-1))/tanh(k)), with dx/dr = 1/r'(x); the reference's own
non-uniform maps (mhdsph_mpi.f90:327-336) are a different stretching of the same kind.
Random numbers come from a counter-based SplitMix64 stream with Box-Muller normals, so any
machine regenerates identical inputs from (seed, index).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib

MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, idx: np.ndarray) -> np.ndarray:
    """SplitMix64 output for counter values idx under a seed (uint64, vectorized)."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + idx.astype(np.uint64) *
             np.uint64(0x9E3779B97F4A7C15) + np.uint64(0x632BE59BD9B4E5B5))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n uniform doubles in (0,1) from the 53 high bits of SplitMix64."""
    z = splitmix64(seed, np.arange(offset, offset + n, dtype=np.uint64))
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def normal(seed: int, n: int) -> np.ndarray:
    """n standard normals (Box-Muller on pairs of uniforms)."""
    m = (n + 1) // 2
    u1 = uniform01(seed, m, 0)
    u2 = uniform01(seed, m, m)
    rad = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * m)
    out[0::2] = rad * np.cos(2 * np.pi * u2)
    out[1::2] = rad * np.sin(2 * np.pi * u2)
    return out[:n]


@dataclass
class Config:
    """One subdomain of the benchmark configurations (reference namelist values made explicit)."""
    nr: int = 128
    nth: int = 256
    nphi: int = 512
    r1: float = 0.72
    r2: float = 0.97
    th1: float = np.pi / 4
    th2: float = 3 * np.pi / 4
    phi1: float = 0.0
    phi2: float = np.pi / 2
    stretch: float = 1.5            # tanh stretching k of the radial grid (0 = uniform)
    gamma: float = 5.0 / 3.0
    Cp: float = 1.0
    rho0: float = 1.0               # density at r2
    T0: float = 1.0                 # temperature at r2 (also the top relaxation target)
    contrast: float = 20.0          # rho(r1)/rho(r2)
    omega0: float = 0.0
    samp: float = 1e-3              # entropy perturbation amplitude
    bamp: float = 1e-4              # vector potential amplitude (x r2)
    seed_s: int = 2
    seed_a: int = 3
    magnetic: bool = True
    xi_cs: float = 1.0
    damping_time: float = 1e9
    cfl_safety: float = 0.4
    SLDiffusion: bool = True        # SLD diffusion folded into RK4 stage 1 (sld_coef 0 gives none)
    sld_coef: float = 0.0
    sld_speedh: float = 1.0
    sld_speedr: float = 0.0
    sld_exp: float = 0.0
    dstype: int = 1                 # SLD speed includes |u|^2 and the Alfven speed^2
    sld_pr: float = 1.0
    sld_prm: float = 1.0
    Mass_Diffusion: bool = True     # diffusion.f90:11 defaults
    Magnetic_Diffusion: bool = True
    Do_Viscous_Heating: bool = True
    Do_Ohmic_Heating: bool = True
    # boundary flags (constants.f90:41-62), the benchmark's walls
    Top_Stress_Free: bool = True
    Bottom_Stress_Free: bool = True
    Top_Constant_Temp: bool = True
    Top_Constant_Flux: bool = False
    Bottom_Constant_Temp: bool = False
    Bottom_Constant_Flux: bool = True
    Top_Radial_Field: bool = True
    Bottom_Radial_Field: bool = True
    Theta_Normal_Field: bool = True
    Theta_Perf_Conductor: bool = False
    Top_Perfect_Cond: bool = False
    Bottom_Perfect_Cond: bool = False
    Theta_Stress_Free: bool = True
    # explicit diffusion (constants.f90:21,66; mhdsph_mpi.f90:349-354): Laplacian = viscosity
    # mur = sld_pr kappa rho_bar, thermal diffusion ks = kappa rho_bar T_bar (rho_bar/rho_bar(r2))^g0,
    # resistivity etar = ks; SLaplacian = the thermal diffusion alone.
    Laplacian: bool = False
    SLaplacian: bool = False
    kappa_value: float = 1e-5
    g0: float = 0.0
    lumr1: float = 0.0              # constant-flux wall luminosities under Laplacian/SLaplacian
    lumr2: float = 0.0

    @property
    def nv(self):
        return 8 if self.magnetic else 5

    @property
    def npts(self):
        return self.nr * self.nth * self.nphi


CONFIGS = {
    "c1": dict(nr=64, nth=64, nphi=128),
    "c2": dict(nr=128, nth=256, nphi=512),
    "c4": dict(nr=256, nth=512, nphi=1024),
    "c5": dict(nr=64, nth=128, nphi=256),
}


def boundary_codes(c: Config):
    """ibcr/ibct (index 1..12) from the boundary flags, mhdsph_mpi.f90:1302-1462."""
    if c.Top_Constant_Temp == c.Top_Constant_Flux:
        raise ValueError("choose one of Top_Constant_Temp / Top_Constant_Flux")
    if c.Bottom_Constant_Temp == c.Bottom_Constant_Flux:
        raise ValueError("choose one of Bottom_Constant_Temp / Bottom_Constant_Flux")
    tsf, bsf = c.Top_Stress_Free, c.Bottom_Stress_Free
    lbcr_vw = {(False, False): 1, (False, True): 2, (True, False): 3, (True, True): 4}[(tsf, bsf)]
    if c.Top_Constant_Temp and c.Bottom_Constant_Temp:
        lbcr_s = 1
    elif c.Top_Constant_Temp and c.Bottom_Constant_Flux:
        lbcr_s = 2
    elif c.Top_Constant_Flux and c.Bottom_Constant_Temp:
        lbcr_s = 3
    else:
        lbcr_s = 4
    lbcr_u = 1
    lbcr_rho = lbcr_t = 4
    lbct_rho, lbct_t, lbct_w = 4, 0, 1
    lbct_u = 0 if c.Theta_Stress_Free else 1
    lbct_v, lbct_s = 4, 0
    ibcr = [0] * 13
    ibct = [0] * 13
    ibcr[1:7] = [lbcr_rho, lbcr_vw, lbcr_vw, lbcr_u, lbcr_s, lbcr_t]
    ibct[1:7] = [lbct_rho, lbct_w, lbct_v, lbct_u, lbct_s, lbct_t]
    if c.magnetic:
        br = bt = bp = ar = at = ap = 1
        if c.Top_Radial_Field and c.Bottom_Radial_Field:
            br, bt, bp, ar, at, ap = 4, 1, 1, 1, 0, 0
        if c.Top_Perfect_Cond and c.Bottom_Perfect_Cond:
            br = bt = bp = ar = at = ap = 1
        tbr = tbt = tbp = tar = tat = tap = 1
        if c.Theta_Normal_Field:
            tbr, tbt, tbp, tar, tat, tap = 1, 4, 1, 0, 1, 0
        if c.Theta_Perf_Conductor:
            tbr = tbt = tbp = tar = tat = tap = 1
        ibcr[7:13] = [bt, bp, br, at, ap, ar]
        ibct[7:13] = [tbt, tbp, tbr, tat, tap, tar]
    return ibcr, ibct


class Model:
    """Grid, background, constants and codes of one subdomain."""

    def __init__(self, c: Config):
        self.c = c
        nr, nth, nphi = c.nr, c.nth, c.nphi
        x = np.linspace(0.0, 1.0, nr)
        self.x = x
        k = c.stretch
        a, b = 0.5 * (c.r1 + c.r2), 0.5 * (c.r2 - c.r1)
        if k > 0:
            tk = np.tanh(k)
            tz = np.tanh(k * (2 * x - 1))
            self.r = a + b * tz / tk
            rp = b * 2 * k * (1 - tz ** 2) / tk                              # dr/dx
            rpp = -8 * k * k * b * (1 - tz ** 2) * tz / tk                    # d2r/dx2
            xm = -1.0 / (nr - 1)
            r_below = a + b * np.tanh(k * (2 * xm - 1)) / tk
        else:
            self.r = a + b * (2 * x - 1)
            rp = np.full(nr, 2 * b)
            rpp = np.zeros(nr)
            r_below = self.r[0] - (self.r[1] - self.r[0])
        self.dxdr = 1.0 / rp
        self.d2xdr2 = -rpp / rp ** 3
        self.dr = np.empty(nr)                     # dr_ind as mhdsph_mpi.f90:331-334
        self.dr[0] = self.r[0] - r_below
        self.dr[1:] = np.diff(self.r)
        self.theta = c.th1 + (c.th2 - c.th1) / (nth - 1) * np.arange(nth)
        self.dphi = (c.phi2 - c.phi1) / nphi
        self.phi = c.phi1 + self.dphi * np.arange(nphi)
        self.dri = float(nr - 1) / (x[-1] - x[0])   # x-space, derivatives.f90:91,95
        self.dti = float(nth - 1) / (self.theta[-1] - self.theta[0])
        self.dpi = float(nphi - 1) / (self.phi[-1] - self.phi[0])
        self.dth = (c.th2 - c.th1) / (nth - 1)
        # constants (rhs_mpi.f90:59-73)
        g = c.gamma
        self.Cv = c.Cp / g
        self.gamm1 = g - 1.0
        self.pconst = self.gamm1 * c.Cp
        self.invtconst = (1.0 / self.Cv) / self.gamm1
        # isentropic polytrope (mhdsph_mpi.f90:317-361) with alp set by the density contrast
        n_poly = 1.0 / self.gamm1
        self.alp = (c.contrast ** (1.0 / n_poly) - 1.0) / (c.r2 / c.r1 - 1.0)
        self.GM = self.alp * c.Cp * c.T0 * c.r2
        zeta = 1.0 + self.alp * (c.r2 / self.r - 1.0)
        self.rho_bar = c.rho0 * zeta ** n_poly
        self.T_bar = c.T0 * zeta
        P2 = self.gamm1 * self.Cv * c.rho0 * c.T0
        self.s_bar = np.full(nr, self.Cv * (np.log(P2) - g * np.log(c.rho0)))
        self.gravity = self.GM / self.r ** 2
        self.eps = np.zeros(nr)
        self.dsdr1 = np.zeros(nth)     # dsdrm = 0 (isentropic), mhdsph_mpi.f90:351,573
        self.dsdr2 = np.zeros(nth)
        self.ibcr, self.ibct = boundary_codes(c)
        # explicit-diffusion profiles (mhdsph_mpi.f90:349-354): exact log-derivatives of the
        # polytrope's rho_bar and T_bar
        self.ks = c.kappa_value * self.rho_bar * self.T_bar * (self.rho_bar / self.rho_bar[-1]) ** c.g0
        self.etar = self.ks.copy()
        self.mur = c.sld_pr * c.kappa_value * self.rho_bar
        dzeta = -self.alp * c.r2 / self.r ** 2 / zeta                          # d ln zeta / dr
        self.dlnks = (g + c.g0) / self.gamm1 * dzeta
        self.dlnmu = dzeta / self.gamm1
        # SLD diffusion speed^2 (diffusion.f90:48-58), flux cutoff off (tanh -> 1)
        ds = c.sld_speedh * (self.rho_bar / self.rho_bar[-1]) ** c.sld_exp + c.sld_speedr
        sp = np.empty(nr + 2)
        sp[1:nr + 1] = ds
        sp[0] = ds[0]
        sp[nr + 1] = c.sld_speedh + c.sld_speedr
        self.diffusion_speed = sp ** 2

    def initial_state(self, device=None):
        """vars0 (nv, nr, nphi, nth): isentropic rho, zero velocity, perturbed entropy, random A.

        device=None builds it with numpy on the host; a torch CUDA device builds the same recipe
        on that GPU (the normals from the same SplitMix64 stream by cssx_normal_fill, equal to
        the host's to rounding) and returns a contiguous FP64 tensor there."""
        if device is not None:
            return self._initial_state_device(device)
        c = self.c
        nr, nth, nphi = c.nr, c.nth, c.nphi
        st = np.zeros((c.nv, nr, nphi, nth))
        st[0] = self.rho_bar[:, None, None]
        xth = np.pi * (self.theta - c.th1) / (c.th2 - c.th1)
        yr = np.pi * (self.r - c.r1) / (c.r2 - c.r1)
        win = (1 - np.cos(yr) ** 16)[:, None, None] * (1 - np.cos(xth) ** 16)[None, None, :]
        xi = normal(c.seed_s, c.npts).reshape(nr, nphi, nth)
        st[4] = self.s_bar[:, None, None] * (1.0 + c.samp * xi * win)
        if c.magnetic:
            wa = (1 - np.cos(yr) ** 2)[:, None, None] * (1 - np.cos(xth) ** 2)[None, None, :]
            for comp in range(3):
                xa = normal(c.seed_a * 1000003 + comp, c.npts).reshape(nr, nphi, nth)
                st[5 + comp] = c.bamp * c.r2 * xa * wa
        return st

    def _initial_state_device(self, device):
        import torch
        c = self.c
        nr, nth, nphi = c.nr, c.nth, c.nphi
        dev = torch.device(device)
        st = torch.zeros((c.nv, nr, nphi, nth), dtype=torch.float64, device=dev)
        stream = ctypes.c_void_p(int(torch.cuda.current_stream(dev).cuda_stream))

        def normals(seed, out):
            _lib.check(_lib.lib().cssx_normal_fill(out.data_ptr(), _lib.PREC_F64, out.numel(), seed, stream),
                       "cssx_normal_fill")
            return out

        def prof(a):
            return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)

        st[0] = prof(self.rho_bar)[:, None, None]
        xth = np.pi * (self.theta - c.th1) / (c.th2 - c.th1)
        yr = np.pi * (self.r - c.r1) / (c.r2 - c.r1)
        win = prof(1 - np.cos(yr) ** 16)[:, None, None] * prof(1 - np.cos(xth) ** 16)[None, None, :]
        xi = normals(c.seed_s, torch.empty((nr, nphi, nth), dtype=torch.float64, device=dev))
        st[4] = prof(self.s_bar)[:, None, None] * (1.0 + c.samp * xi * win)
        del xi
        if c.magnetic:
            wa = prof(1 - np.cos(yr) ** 2)[:, None, None] * prof(1 - np.cos(xth) ** 2)[None, None, :]
            for comp in range(3):
                a = normals(c.seed_a * 1000003 + comp, st[5 + comp])
                a.mul_(c.bamp * c.r2).mul_(wa)
        return st

    # ---- C-ABI structs -------------------------------------------------------------------------
    def grid_struct(self):
        return _lib.Grid(self.c.nr, self.c.nth, self.c.nphi, 1, 1, 1, 1, self.dri, self.dti, self.dpi)

    def params_struct(self):
        c = self.c
        p = MhdParams()
        p.gamma, p.Cp, p.omega0, p.xi_cs = c.gamma, c.Cp, c.omega0, c.xi_cs
        p.damping_time, p.T0 = c.damping_time, c.T0
        p.magnetic = int(c.magnetic)
        p.top_relax = 1
        p.bottom_stress_free, p.top_stress_free = int(c.Bottom_Stress_Free), int(c.Top_Stress_Free)
        p.bottom_const_flux, p.top_const_flux = int(c.Bottom_Constant_Flux), int(c.Top_Constant_Flux)
        p.bottom_const_temp, p.top_const_temp = int(c.Bottom_Constant_Temp), int(c.Top_Constant_Temp)
        p.bottom_radial_field, p.top_radial_field = int(c.Bottom_Radial_Field), int(c.Top_Radial_Field)
        p.theta_normal_field, p.theta_perf_cond = int(c.Theta_Normal_Field), int(c.Theta_Perf_Conductor)
        p.theta_stress_free = int(c.Theta_Stress_Free)
        for i in range(13):
            p.ibcr[i] = self.ibcr[i]
            p.ibct[i] = self.ibct[i]
        p.sld_on = int(c.SLDiffusion)
        p.sld_coef = c.sld_coef
        p.mass_diffusion, p.magnetic_diffusion = int(c.Mass_Diffusion), int(c.Magnetic_Diffusion)
        p.viscous_heating, p.ohmic_heating = int(c.Do_Viscous_Heating), int(c.Do_Ohmic_Heating)
        p.dstype = c.dstype
        p.sld_pr, p.sld_prm = c.sld_pr, c.sld_prm
        p.bottom_perf_cond, p.top_perf_cond = int(c.Bottom_Perfect_Cond), int(c.Top_Perfect_Cond)
        p.laplacian, p.slaplacian = int(c.Laplacian), int(c.SLaplacian)
        p.lumr1, p.lumr2 = c.lumr1, c.lumr2
        return p

    def profiles_struct(self):
        """cssx_mhd_profiles pointing at this model's arrays (keep the Model alive)."""
        self._keep = {k: np.ascontiguousarray(getattr(self, k), dtype=np.float64) for k in
                      ("r", "dr", "dxdr", "theta", "gravity", "eps", "rho_bar", "s_bar", "dsdr1", "dsdr2",
                       "diffusion_speed", "mur", "dlnmu", "ks", "dlnks", "etar", "d2xdr2")}
        sr1 = np.full(self.c.nth, self.s_bar[0])
        sr2 = np.full(self.c.nth, self.s_bar[-1])
        self._keep["sr1"], self._keep["sr2"] = sr1, sr2
        pr = MhdProfiles()
        for k, v in self._keep.items():
            setattr(pr, k, v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return pr


_dp = ctypes.POINTER(ctypes.c_double)


class MhdParams(ctypes.Structure):
    """cssx_mhd_params (include/cssx.h)."""
    _fields_ = [("gamma", ctypes.c_double), ("Cp", ctypes.c_double), ("omega0", ctypes.c_double),
                ("xi_cs", ctypes.c_double), ("damping_time", ctypes.c_double), ("T0", ctypes.c_double),
                ("magnetic", ctypes.c_int), ("top_relax", ctypes.c_int),
                ("bottom_stress_free", ctypes.c_int), ("top_stress_free", ctypes.c_int),
                ("bottom_const_flux", ctypes.c_int), ("top_const_flux", ctypes.c_int),
                ("bottom_const_temp", ctypes.c_int), ("top_const_temp", ctypes.c_int),
                ("bottom_radial_field", ctypes.c_int), ("top_radial_field", ctypes.c_int),
                ("theta_normal_field", ctypes.c_int), ("theta_perf_cond", ctypes.c_int),
                ("theta_stress_free", ctypes.c_int),
                ("ibcr", ctypes.c_int * 13), ("ibct", ctypes.c_int * 13),
                ("sld_on", ctypes.c_int), ("sld_coef", ctypes.c_double),
                ("mass_diffusion", ctypes.c_int), ("magnetic_diffusion", ctypes.c_int),
                ("viscous_heating", ctypes.c_int), ("ohmic_heating", ctypes.c_int),
                ("dstype", ctypes.c_int), ("sld_pr", ctypes.c_double), ("sld_prm", ctypes.c_double),
                ("bottom_perf_cond", ctypes.c_int), ("top_perf_cond", ctypes.c_int),
                ("laplacian", ctypes.c_int), ("slaplacian", ctypes.c_int),
                ("lumr1", ctypes.c_double), ("lumr2", ctypes.c_double)]


class MhdProfiles(ctypes.Structure):
    """cssx_mhd_profiles (include/cssx.h)."""
    _fields_ = [(k, _dp) for k in ("r", "dr", "dxdr", "theta", "gravity", "eps", "rho_bar", "s_bar",
                                   "dsdr1", "dsdr2", "sr1", "sr2", "diffusion_speed", "mur", "dlnmu", "ks",
                                   "dlnks", "etar", "d2xdr2")]
