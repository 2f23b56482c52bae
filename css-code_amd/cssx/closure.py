"""Learned subgrid closure: a coarse grid standing in for a fine one.

The reference has no learned model; its subgrid model is the SLD diffusion of diffusion.f90.
This module adds, beside it, a small network that predicts what a coarse grid's RHS misses:

  * training data come from a fine run of the same recipe: at snapshots, the fine state and
    the fine RHS are box-filtered by `factor` onto the coarse grid; the target is the filtered
    fine RHS minus the coarse RHS of the filtered state (8 fields);
  * inputs are the 3x3x3 stencils of the 8 filtered fields (216 features, rho and s as
    fluctuations about the hydrostatic background), standardized;
  * the network 216 -> 128 -> 128 -> 8 (GELU) is trained with Adam on the GPU by the library's
    own BF16 matrix-core forward/backward kernels (FP32 master weights and accumulation);
  * a coarse run attaches it (cssx_mhd_set_closure): every RHS evaluation adds its tendency; the
    plain coarse run (SLD only) is the fallback it is compared with, and the filtered fine run
    is the reference for both.

Everything heavy runs in libcssx.so (csrc/cssx_closure.hip); torch holds the device memory.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .mhd import Solver, _stream
from .model import Config, Model

NFEAT = 216
NOUT = 8
FIELDS = ("rho", "w", "v", "u", "s", "A_theta", "A_phi", "A_r")


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _cuda(x, dtype=None):
    if not (isinstance(x, torch.Tensor) and x.is_cuda and x.is_contiguous()):
        raise ValueError("a contiguous CUDA tensor is required")
    if dtype is not None and x.dtype != dtype:
        raise TypeError(f"{dtype} tensor required, got {x.dtype}")
    return x


def _prec(x):
    if x.dtype == torch.float64:
        return _lib.PREC_F64
    if x.dtype == torch.float32:
        return _lib.PREC_F32
    raise TypeError("float32 or float64 tensor expected")


class Network:
    """The closure network (cssx_mlp): 216 -> h1 -> h2 -> 8, GELU, BF16 matrix cores."""

    def __init__(self, h1: int = 128, h2: int = 128, max_rows: int = 262144, seed: int = 7):
        if not torch.cuda.is_available():
            raise RuntimeError("cssx.closure needs a ROCm GPU")
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().cssx_mlp_create(NFEAT, h1, h2, NOUT, max_rows, seed, ctypes.byref(h)),
                   "cssx_mlp_create")
        self._h = h
        self.h1, self.h2, self.max_rows = h1, h2, max_rows
        self.n_params = int(_lib.lib().cssx_mlp_num_params(h))
        self.norm = None
        self._loss = torch.zeros(1, dtype=torch.float32, device="cuda")

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().cssx_mlp_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        p = np.empty(self.n_params, dtype=np.float32)
        _lib.check(_lib.lib().cssx_mlp_get_params(self._h, _fptr(p), None), "cssx_mlp_get_params")
        return p

    def grads(self) -> np.ndarray:
        g = np.empty(self.n_params, dtype=np.float32)
        _lib.check(_lib.lib().cssx_mlp_get_params(self._h, None, _fptr(g)), "cssx_mlp_get_params")
        return g

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n_params:
            raise ValueError(f"{self.n_params} parameters expected")
        _lib.check(_lib.lib().cssx_mlp_set_params(self._h, _fptr(p)), "cssx_mlp_set_params")

    def set_norm(self, x_mean, x_std, y_mean, y_std) -> None:
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (x_mean, x_std, y_mean, y_std)]
        if arrs[0].size != NFEAT or arrs[1].size != NFEAT or arrs[2].size != NOUT or arrs[3].size != NOUT:
            raise ValueError("norm sizes: 216, 216, 8, 8")
        _lib.check(_lib.lib().cssx_mlp_set_norm(self._h, *[_fptr(a) for a in arrs]), "cssx_mlp_set_norm")
        self.norm = tuple(arrs)

    def forward(self, xb: torch.Tensor) -> torch.Tensor:
        """Standardized outputs (rows, 8) FP32 of standardized BF16 inputs (rows, 216)."""
        _cuda(xb, torch.bfloat16)
        rows = xb.shape[0]
        y = torch.empty((rows, NOUT), dtype=torch.float32, device=xb.device)
        _lib.check(_lib.lib().cssx_mlp_forward(self._h, xb.data_ptr(), rows, y.data_ptr(), _stream()),
                   "cssx_mlp_forward")
        return y

    def grad(self, xb: torch.Tensor, t: torch.Tensor) -> torch.Tensor:
        """Gradient of the mean squared error into the network (read with grads()); returns
        the loss as a device tensor."""
        _cuda(xb, torch.bfloat16)
        _cuda(t, torch.float32)
        _lib.check(_lib.lib().cssx_mlp_grad(self._h, xb.data_ptr(), t.data_ptr(), xb.shape[0],
                                            self._loss.data_ptr(), _stream()), "cssx_mlp_grad")
        return self._loss

    def train_step(self, xb: torch.Tensor, t: torch.Tensor, lr: float, loss_out: torch.Tensor | None = None):
        _cuda(xb, torch.bfloat16)
        _cuda(t, torch.float32)
        lo = self._loss if loss_out is None else loss_out
        _lib.check(_lib.lib().cssx_mlp_train_step(self._h, xb.data_ptr(), t.data_ptr(), xb.shape[0], float(lr),
                                                  lo.data_ptr(), _stream()), "cssx_mlp_train_step")
        return lo


# ---- data kernels -------------------------------------------------------------------------------
def box_filter(x: torch.Tensor, factor: int) -> torch.Tensor:
    """(nv, nr, nphi, nth) -> FP32 (nv, nr/f, nphi/f, nth/f), mean over factor^3 blocks."""
    _cuda(x)
    nv, nr, nphi, nth = x.shape
    out = torch.empty((nv, nr // factor, nphi // factor, nth // factor), dtype=torch.float32, device=x.device)
    _lib.check(_lib.lib().cssx_box_filter(x.data_ptr(), _prec(x), nv, nr, nphi, nth, factor, out.data_ptr(),
                                          _stream()), "cssx_box_filter")
    return out


def _background(model: Model, dtype, device="cuda"):
    """rho_bar, s_bar on the device, uploaded once per (model, dtype, device): a closure
    tendency runs once per RHS evaluation, and a host-to-device copy per call would stall it."""
    cache = model.__dict__.setdefault("_closure_background", {})
    key = (dtype, str(device))
    if key not in cache:
        cache[key] = (torch.as_tensor(model.rho_bar, dtype=dtype, device=device).contiguous(),
                      torch.as_tensor(model.s_bar, dtype=dtype, device=device).contiguous())
    return cache[key]


def features(state: torch.Tensor, model: Model, fluct: bool = False, net: Network | None = None,
             row0: int = 0, rows: int | None = None) -> torch.Tensor:
    """Features of rows [row0, row0+rows) of an 8-field state: standardized BF16 with `net`
    (its norm), raw FP32 without."""
    _cuda(state)
    _, nr, nphi, nth = state.shape
    N = nr * nphi * nth
    rows = N - row0 if rows is None else rows
    rb, sb = _background(model, state.dtype, state.device)
    if net is not None:
        out = torch.empty((rows, NFEAT), dtype=torch.bfloat16, device=state.device)
        args = (out.data_ptr(), None)
    else:
        out = torch.empty((rows, NFEAT), dtype=torch.float32, device=state.device)
        args = (None, out.data_ptr())
    _lib.check(_lib.lib().cssx_closure_features(net.handle if net else None, state.data_ptr(), _prec(state),
                                                int(fluct), rb.data_ptr(), sb.data_ptr(), nr, nphi, nth, row0,
                                                rows, *args, _stream()), "cssx_closure_features")
    return out


def tendency(net: Network, state: torch.Tensor, model: Model, fluct: bool = False,
             out: torch.Tensor | None = None) -> torch.Tensor:
    """The network's closure tendency (8, nr, nphi, nth) FP32 of an 8-field state."""
    _cuda(state)
    _, nr, nphi, nth = state.shape
    rb, sb = _background(model, state.dtype, state.device)
    if out is None:
        out = torch.empty(state.shape, dtype=torch.float32, device=state.device)
    _lib.check(_lib.lib().cssx_closure_tendency(net.handle, state.data_ptr(), _prec(state), int(fluct), rb.data_ptr(),
                                                sb.data_ptr(), nr, nphi, nth, out.data_ptr(), _stream()),
               "cssx_closure_tendency")
    return out


def residual_rows(fine_filtered: torch.Tensor, coarse_rhs: torch.Tensor) -> torch.Tensor:
    """Row-major (N, 8) FP32 filtered-fine minus coarse tendency."""
    _cuda(fine_filtered, torch.float32)
    _cuda(coarse_rhs)
    N = fine_filtered[0].numel()
    raw = torch.empty((N, NOUT), dtype=torch.float32, device=fine_filtered.device)
    _lib.check(_lib.lib().cssx_closure_targets(fine_filtered.data_ptr(), coarse_rhs.data_ptr(), _prec(coarse_rhs),
                                               N, None, None, None, raw.data_ptr(), _stream()),
               "cssx_closure_targets")
    return raw


def colstats(x: torch.Tensor):
    """Column mean and standard deviation of a row-major FP32 matrix (synchronizes)."""
    _cuda(x, torch.float32)
    rows, cols = x.shape
    m = np.empty(cols)
    s = np.empty(cols)
    _lib.check(_lib.lib().cssx_colstats(x.data_ptr(), rows, cols, _dptr(m), _dptr(s), _stream()), "cssx_colstats")
    return m, s


def standardize(x: torch.Tensor, mean, std, to_bf16: bool) -> torch.Tensor:
    _cuda(x, torch.float32)
    rows, cols = x.shape
    m = torch.as_tensor(np.asarray(mean, np.float32), device=x.device)
    s = torch.as_tensor(np.asarray(std, np.float32), device=x.device)
    out = torch.empty((rows, cols), dtype=torch.bfloat16 if to_bf16 else torch.float32, device=x.device)
    _lib.check(_lib.lib().cssx_standardize_rows(x.data_ptr(), rows, cols, m.data_ptr(), s.data_ptr(),
                                                out.data_ptr() if to_bf16 else None,
                                                None if to_bf16 else out.data_ptr(), _stream()),
               "cssx_standardize_rows")
    return out


def gather_rows(src: torch.Tensor, idx: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
    _cuda(src)
    _cuda(idx, torch.int64)
    n = idx.numel()
    if out is None:
        out = torch.empty((n,) + tuple(src.shape[1:]), dtype=src.dtype, device=src.device)
    row_bytes = src[0].numel() * src.element_size()
    _lib.check(_lib.lib().cssx_gather_rows(src.data_ptr(), row_bytes, idx.data_ptr(), n, out.data_ptr(), _stream()),
               "cssx_gather_rows")
    return out


def field_stats(state: torch.Tensor, model: Model, fluct: bool = False):
    """Per-field mean and RMS (rho and s about the background unless fluct); synchronizes."""
    _cuda(state)
    nv, nr, nphi, nth = state.shape
    rb, sb = _background(model, state.dtype, state.device)
    m = np.empty(nv)
    r = np.empty(nv)
    _lib.check(_lib.lib().cssx_field_stats(state.data_ptr(), _prec(state), nv, nr, nphi, nth, int(fluct),
                                           rb.data_ptr(), sb.data_ptr(), _dptr(m), _dptr(r), _stream()),
               "cssx_field_stats")
    return m, r


# ---- training data ------------------------------------------------------------------------------
@dataclass
class Snapshot:
    x_raw: torch.Tensor     # (Nc, 216) FP32 raw features of the filtered state
    r_raw: torch.Tensor     # (Nc, 8) FP32 filtered-fine minus coarse tendency
    time: float


def coarse_config(fine: Config, factor: int) -> Config:
    d = dict(fine.__dict__)
    d.update(nr=fine.nr // factor, nth=fine.nth // factor, nphi=fine.nphi // factor)
    return Config(**d)


def filter_state(y: torch.Tensor, fine: Solver, coarse: Solver, factor: int) -> torch.Tensor:
    """The coarse FP64 state of a fine FP64 state: the box-filtered fluctuations (rho and s about
    the fine hydrostatic background, the other fields as they are) on top of the coarse grid's
    own background.  Filtering the full fields instead would hand the coarse grid a density
    profile out of its discrete hydrostatic balance (the filter averages the stratification over
    a cell, and the coarse points are not the filter's cell centres), whose spurious buoyancy
    would swamp every subgrid effect."""
    yf = box_filter(y - fine.background(y.device), factor).double()
    return (yf + coarse.background(y.device)).contiguous()


def snapshot(fine: Solver, coarse: Solver, y: torch.Tensor, factor: int, t: float,
             include_sld: bool = False) -> Snapshot:
    """Filtered fine tendency minus coarse tendency of the filtered FP64 state y (one time): the
    RHS, plus with include_sld the SLD diffusion each grid applies (the coarse grid's SLD acts
    on scales the fine grid resolves, so its surplus damping is part of what the closure can
    learn to undo)."""
    F = fine.rhs(y)
    if include_sld:
        F += fine.diffusion(y)
    Fb = box_filter(F, factor)
    del F
    Uc = filter_state(y, fine, coarse, factor)
    Rc = coarse.rhs(Uc)
    if include_sld:
        Rc += coarse.diffusion(Uc)
    return Snapshot(features(Uc, coarse.model), residual_rows(Fb, Rc), t)


@dataclass
class TrainingSet:
    X: torch.Tensor         # (n, 216) BF16 standardized
    T: torch.Tensor         # (n, 8) FP32 standardized
    x_mean: np.ndarray
    x_std: np.ndarray
    y_mean: np.ndarray
    y_std: np.ndarray


def build_training_set(snaps, samples: int, seed: int = 7) -> TrainingSet:
    """`samples` rows drawn without replacement, evenly over the snapshots."""
    rng = np.random.default_rng(seed)
    per = [samples // len(snaps) + (1 if i < samples % len(snaps) else 0) for i in range(len(snaps))]
    xs, rs = [], []
    for s, k in zip(snaps, per):
        n = s.x_raw.shape[0]
        idx = torch.as_tensor(np.sort(rng.choice(n, size=min(k, n), replace=False)), device=s.x_raw.device)
        xs.append(gather_rows(s.x_raw, idx))
        rs.append(gather_rows(s.r_raw, idx))
    Xr = torch.cat(xs)
    Rr = torch.cat(rs)
    xm, xsd = colstats(Xr)
    ym, ysd = colstats(Rr)
    xsd = np.where(xsd > 0, xsd, 1.0)
    ysd = np.where(ysd > 0, ysd, 1.0)
    return TrainingSet(standardize(Xr, xm, xsd, True), standardize(Rr, ym, ysd, False), xm, xsd, ym, ysd)


def train(net: Network, ts: TrainingSet, epochs: int, batch: int = 8192, lr: float = 1e-3, seed: int = 7):
    """Adam over shuffled minibatches; returns the mean training loss of every epoch."""
    net.set_norm(ts.x_mean, ts.x_std, ts.y_mean, ts.y_std)
    n = ts.X.shape[0]
    nb = n // batch
    if nb < 1:
        raise ValueError(f"need at least one batch of {batch} rows (have {n})")
    rng = np.random.default_rng(seed)
    xb = torch.empty((batch, NFEAT), dtype=torch.bfloat16, device=ts.X.device)
    tb = torch.empty((batch, NOUT), dtype=torch.float32, device=ts.X.device)
    losses = torch.zeros(nb, dtype=torch.float32, device=ts.X.device)
    hist = []
    for _ in range(epochs):
        perm = torch.as_tensor(rng.permutation(n)[: nb * batch], device=ts.X.device)
        for b in range(nb):
            idx = perm[b * batch:(b + 1) * batch]
            gather_rows(ts.X, idx, xb)
            gather_rows(ts.T, idx, tb)
            net.train_step(xb, tb, lr, losses[b:b + 1])
        hist.append(float(losses.mean().item()))
    return hist


def predict(net: Network, x_raw: torch.Tensor, ts: TrainingSet) -> torch.Tensor:
    """Un-standardized network tendency (rows, 8) of raw feature rows (training-set scaling)."""
    xb = standardize(x_raw, ts.x_mean, ts.x_std, True)
    n = xb.shape[0]
    pred = torch.empty((n, NOUT), dtype=torch.float32, device=xb.device)
    for r0 in range(0, n, net.max_rows):
        r1 = min(n, r0 + net.max_rows)
        pred[r0:r1] = net.forward(xb[r0:r1])
    ym = torch.as_tensor(ts.y_mean, dtype=torch.float32, device=xb.device)
    ys = torch.as_tensor(ts.y_std, dtype=torch.float32, device=xb.device)
    return pred * ys + ym


def apriori_gain(net: Network, snap: Snapshot, ts: TrainingSet, mask=None):
    """Per-field RMS of a snapshot's residual without and with the network's correction (fields
    with mask False keep the plain residual)."""
    p = predict(net, snap.x_raw, ts)
    if mask is not None:
        p = p * torch.as_tensor(np.asarray(mask, np.float32), device=p.device)
    before = snap.r_raw.pow(2).mean(0).sqrt().cpu().numpy()
    after = (snap.r_raw - p).pow(2).mean(0).sqrt().cpu().numpy()
    return before, after


def field_gate(net: Network, validation: Snapshot, ts: TrainingSet, threshold: float = 1.0):
    """Fields whose residual RMS the network brings below `threshold` times its plain value on a
    validation snapshot; the others fall back to the plain coarse RHS (their correction is
    switched off in the network's norm)."""
    before, after = apriori_gain(net, validation, ts)
    mask = after < threshold * before
    net.set_norm(ts.x_mean, ts.x_std, np.where(mask, ts.y_mean, 0.0), np.where(mask, ts.y_std, 0.0))
    return mask


# ---- coarse runs ----------------------------------------------------------------------------------
def kinetic_energy(state: torch.Tensor) -> float:
    """Volume (index-space) mean of rho |u|^2 / 2 of an FP64 state."""
    return float((0.5 * state[0] * (state[1] ** 2 + state[2] ** 2 + state[3] ** 2)).mean().item())


SCALARS = ("kinener", "cke", "mcke", "drke", "magener")


def stats_of(state: torch.Tensor, sol: Solver) -> dict:
    """Per-field RMS (rho and s about the background) and the Scalar_Output energies of
    report_mpi.f90:173-262 (kinetic energy, its radial / radial+meridional / azimuthal parts,
    magnetic energy) of an FP64 state on the solver's grid."""
    from .diagnostics import scalar_quantities
    m, r = field_stats(state, sol.model)
    q = scalar_quantities(sol, state)
    out = {"rms": r.tolist(), "ke": kinetic_energy(state)}
    out.update({k: q[k] for k in SCALARS if k in q})
    return out


def compare(fine_cfg: Config, factor: int, net: Network | None, coarse_steps: int, every: int,
            precision: str = "f32", y_fine: torch.Tensor | None = None, progress=None):
    """Runs the fine grid and, from its filtered initial state, the coarse grid with and without
    the closure for `coarse_steps` steps of 4x the fine step (factor fine steps each); every
    `every` coarse steps it records the statistics of the filtered fine state and of both coarse
    states.  Returns the records and the wall times (fine run + filtering; coarse runs).
    progress, if given, is called with each record as it is taken."""
    fine = Solver(Model(fine_cfg))
    cc = coarse_config(fine_cfg, factor)
    cmodel = Model(cc)
    plain = Solver(cmodel)
    closed = Solver(cmodel) if net is not None else None
    if closed is not None:
        _lib.check(_lib.lib().cssx_mhd_set_closure(closed._h, net.handle), "cssx_mhd_set_closure")
    y = y_fine if y_fine is not None else fine.model.initial_state(device="cuda")
    dtf = fine.cfl(y)
    yc0 = filter_state(y, fine, plain, factor)
    dtc = min(factor * dtf, plain.cfl(yc0))
    dtf = dtc / factor
    yp = yc0.clone()
    yq = yc0.clone() if closed is not None else None
    recs = []
    t_fine = t_plain = t_closed = 0.0

    def rec(k):
        yf = filter_state(y, fine, plain, factor)
        r = {"step": k, "time": k * dtc, "filtered_fine": stats_of(yf, plain), "coarse": stats_of(yp, plain)}
        if yq is not None:
            r["coarse_closure"] = stats_of(yq, plain)
        recs.append(r)
        if progress is not None:
            progress(r)

    rec(0)
    for k0 in range(0, coarse_steps, every):
        n = min(every, coarse_steps - k0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fine.rk4(y, dtf, n * factor, precision)
        filter_state(y, fine, plain, factor)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        plain.rk4(yp, dtc, n, precision)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if closed is not None:
            closed.rk4(yq, dtc, n, precision)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        t_fine += t1 - t0
        t_plain += t2 - t1
        t_closed += t3 - t2
        rec(k0 + n)
    fine.close()
    plain.close()
    if closed is not None:
        _lib.check(_lib.lib().cssx_mhd_set_closure(closed._h, None), "cssx_mhd_set_closure")
        closed.close()
    return recs, {"fine_s": t_fine, "coarse_plain_s": t_plain, "coarse_closure_s": t_closed, "dt_coarse": dtc,
                  "dt_fine": dtf}


def calibrate_gain(fine_cfg: Config, factor: int, net: Network, ts: TrainingSet, mask, gains, coarse_steps: int,
                   every: int, precision: str = "f32", progress=None):
    """A-posteriori gain of the gated correction: the coarse grid runs with the correction
    scaled by each gain in `gains` (0 = the plain fallback) from the filtered initial state of
    `fine_cfg` (a calibration case apart from the evaluated one), beside the fine run filtered
    every `every` coarse steps.  An a-priori gain on snapshots does not survive the feedback of
    a run (the corrected field drifts), so the gain is set by the run itself.  Each statistic's
    error is its largest relative error over the records; a gain is scored by the mean of those
    over all statistics (the largest alone is often a statistic no gain changes, such as the
    magnetic energy when A is not corrected).  Returns the chosen gain, each gain's score and
    each gain's largest error, and leaves the network at the chosen gain."""
    fine = Solver(Model(fine_cfg))
    cmodel = Model(coarse_config(fine_cfg, factor))
    plain = Solver(cmodel)
    closed = Solver(cmodel)
    y = fine.model.initial_state(device="cuda")
    dtf = fine.cfl(y)
    yc0 = filter_state(y, fine, plain, factor)
    dtc = min(factor * dtf, plain.cfl(yc0))
    dtf = dtc / factor
    fine_recs = []
    for k0 in range(0, coarse_steps, every):
        n = min(every, coarse_steps - k0)
        fine.rk4(y, dtf, n * factor, precision)
        fine_recs.append(stats_of(filter_state(y, fine, plain, factor), plain))
    fine.close()
    del y
    mask = np.asarray(mask, bool)
    score, worst = [], []
    _lib.check(_lib.lib().cssx_mhd_set_closure(closed._h, net.handle), "cssx_mhd_set_closure")
    for g in gains:
        net.set_norm(ts.x_mean, ts.x_std, np.where(mask, g * ts.y_mean, 0.0), np.where(mask, g * ts.y_std, 0.0))
        yq = yc0.clone()
        err = None   # per statistic, the largest relative error over the records
        for i, k0 in enumerate(range(0, coarse_steps, every)):
            n = min(every, coarse_steps - k0)
            closed.rk4(yq, dtc, n, precision)
            if not torch.isfinite(yq).all():
                err = np.full(1, np.inf)
                break
            a, b = stats_of(yq, plain), fine_recs[i]
            keys = ("ke",) + tuple(k for k in SCALARS if k in b)
            e = np.concatenate([np.abs(np.asarray(a["rms"]) - np.asarray(b["rms"])) /
                                np.maximum(np.abs(np.asarray(b["rms"])), 1e-300),
                                [abs(a[k] - b[k]) / max(abs(b[k]), 1e-300) for k in keys]])
            err = e if err is None else np.maximum(err, e)
        score.append(float(np.mean(err)))
        worst.append(float(np.max(err)))
        if progress is not None:
            progress(g, score[-1], worst[-1])
    _lib.check(_lib.lib().cssx_mhd_set_closure(closed._h, None), "cssx_mhd_set_closure")
    closed.close()
    plain.close()
    best = gains[int(np.argmin(score))]
    net.set_norm(ts.x_mean, ts.x_std, np.where(mask, best * ts.y_mean, 0.0), np.where(mask, best * ts.y_std, 0.0))
    return best, score, worst


def stat_errors(recs, key: str):
    """Max over records of the relative error of each field's RMS and of the kinetic energy
    against the filtered fine run (records after the initial one)."""
    fld = np.zeros(NOUT)
    ke = 0.0
    for r in recs[1:]:
        a = np.asarray(r[key]["rms"])
        b = np.asarray(r["filtered_fine"]["rms"])
        fld = np.maximum(fld, np.abs(a - b) / np.maximum(np.abs(b), 1e-300))
        kb = r["filtered_fine"]["ke"]
        ke = max(ke, abs(r[key]["ke"] - kb) / max(abs(kb), 1e-300))
    return fld, ke


def scalar_errors(recs, key: str) -> dict:
    """Max over records (after the initial one) of the relative error of each Scalar_Output
    energy against the filtered fine run."""
    out = {}
    for name in SCALARS:
        e = 0.0
        for r in recs[1:]:
            if name not in r[key]:
                continue
            b = r["filtered_fine"][name]
            e = max(e, abs(r[key][name] - b) / max(abs(b), 1e-300))
        out[name] = e
    return out
