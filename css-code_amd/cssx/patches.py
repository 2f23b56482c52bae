"""Batched derivatives of many uneven patches (cssx_patchset) and the ensemble of small MHD runs.

A patch is an independent subdomain f(nth, nphi, nr) with walls on all six faces; all three
first derivatives of every patch of a set run as one persistent launch (csrc/cssx_patches.hip).
`patch_shapes` draws the lengths (uniform integers in [lo, hi] per axis from the project's
SplitMix64 generator), `PatchChunk` owns the device memory of a set of patches (one arena per
role, every patch 256-byte aligned), fills it with seeded smoothed noise and runs / autotunes
the launch.

`Ensemble` steps many small MHD subdomains (one Solver each, their own SLD coefficient) and can
capture a whole ensemble step in one HIP graph (torch.cuda.CUDAGraph holds it), so that the
thousands of small launches of a step replay without host launch overhead.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .mhd import Solver, _stream
from .model import Config, Model, uniform01

ALIGN = 64   # elements (256 B in FP32)


def patch_shapes(count: int, seed: int = 4, lo: int = 32, hi: int = 512):
    """(count, 3) integer (nr, nth, nphi) lengths, uniform in [lo, hi]."""
    u = uniform01(seed, 3 * count)
    return (lo + np.floor(u * (hi - lo + 1)).astype(np.int64)).clip(lo, hi).reshape(count, 3)


class PatchChunk:
    """Device buffers, patch table and persistent-launch plan for a list of patch shapes."""

    def __init__(self, shapes, prec: str = "f32", bdry: int = 1, axes=(True, True, True), device="cuda"):
        if not torch.cuda.is_available():
            raise RuntimeError("cssx.patches needs a ROCm GPU")
        self.dtype = torch.float32 if prec == "f32" else torch.float64
        self.prec = _lib.PREC_F32 if prec == "f32" else _lib.PREC_F64
        self.shapes = [tuple(int(x) for x in s) for s in shapes]
        sizes = [s[0] * s[1] * s[2] for s in self.shapes]
        self.offsets = []
        off = 0
        for n in sizes:
            self.offsets.append(off)
            off += (n + ALIGN - 1) // ALIGN * ALIGN
        self.total = off
        self.points = int(sum(sizes))
        self.inp = torch.empty(off, dtype=self.dtype, device=device)
        self.out = [torch.empty(off, dtype=self.dtype, device=device) if a else None for a in axes]
        arr = (_lib.Patch * len(self.shapes))()
        es = self.inp.element_size()
        for i, ((nr, nth, nphi), o) in enumerate(zip(self.shapes, self.offsets)):
            p = arr[i]
            p.nr, p.nth, p.nphi = nr, nth, nphi
            p.in_ = self.inp.data_ptr() + o * es
            p.d_r, p.d_th, p.d_ph = [(t.data_ptr() + o * es) if t is not None else None for t in self.out]
            p.bdry = bdry
            # unit domain per axis: 1/dx = n-1
            p.dri, p.dti, p.dpi = float(nr - 1), float(nth - 1), float(nphi - 1)
        self._arr = arr
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().cssx_patchset_create(arr, len(self.shapes), self.prec, ctypes.byref(h)),
                   "cssx_patchset_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().cssx_patchset_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def field(self, i, which=None):
        """Patch i's input (which=None) or derivative along axis which (0 r, 1 theta, 2 phi),
        as an (nr, nphi, nth) view."""
        nr, nth, nphi = self.shapes[i]
        buf = self.inp if which is None else self.out[which]
        o = self.offsets[i]
        return buf[o:o + nr * nth * nphi].view(nr, nphi, nth)

    def fill(self, seed: int):
        """Every patch: seeded N(0,1) smoothed by [1,2,1]/4 per axis (patch i uses seed*1000003+i)."""
        tmp = torch.empty(max(s[0] * s[1] * s[2] for s in self.shapes), dtype=self.dtype, device=self.inp.device)
        for i, (nr, nth, nphi) in enumerate(self.shapes):
            f = self.field(i)
            _lib.check(_lib.lib().cssx_patch_fill(f.data_ptr(), tmp.data_ptr(), self.prec, nr, nth, nphi,
                                                  seed * 1000003 + i, _stream()), "cssx_patch_fill")

    def run(self, variant: int = -1):
        _lib.check(_lib.lib().cssx_patchset_run(self._h, int(variant), _stream()), "cssx_patchset_run")

    def autotune(self, reps: int = 2):
        best = ctypes.c_int()
        ms = (ctypes.c_float * 4)()
        _lib.check(_lib.lib().cssx_patchset_autotune(self._h, reps, _stream(), ctypes.byref(best), ms),
                   "cssx_patchset_autotune")
        return best.value, list(ms)


def ensemble_sld(ra: np.ndarray, base: float = 0.1, ra0: float = 1e5) -> np.ndarray:
    """SLD coefficient of each member: base * (Ra/Ra0)^(-1/2), the diffusivity scaling of a fixed
    Prandtl number (the SLD coefficient stands in for the explicit diffusivities that Ra sets)."""
    return base * (np.asarray(ra) / ra0) ** -0.5


class Ensemble:
    """`members` small MHD subdomains of one shape, Rayleigh numbers log-uniform in [ra_lo, ra_hi]
    (SplitMix64 seed), each with its own Solver, state and CFL step.

    A member's sweeps launch fewer workgroups than the GPU has CUs (a 48x96x96 r sweep: 144), so
    the members are dealt round-robin over `streams` side streams (default priority) that run
    concurrently; each member owns its scratch, so members on different streams share nothing
    they write."""

    def __init__(self, members: int = 64, shape=(48, 96, 96), seed: int = 4, ra_lo: float = 1e4,
                 ra_hi: float = 1e6, precision: str = "f32", streams: int = 4):
        nr, nth, nphi = shape
        u = uniform01(seed * 7919 + 1, members)
        self.ra = np.exp(np.log(ra_lo) + u * (np.log(ra_hi) - np.log(ra_lo)))
        self.sld = ensemble_sld(self.ra)
        self.precision = precision
        self.solvers, self.states, self.dts = [], [], []
        for k in range(members):
            c = Config(nr=nr, nth=nth, nphi=nphi, sld_coef=float(self.sld[k]), seed_s=seed * 1000 + k,
                       seed_a=seed * 1000 + members + k)
            sol = Solver(Model(c))
            y = sol.model.initial_state(device="cuda")
            self.solvers.append(sol)
            self.states.append(y)
            self.dts.append(sol.cfl(y))
        self.points = members * nr * nth * nphi
        self.graph = None
        self.streams = [torch.cuda.Stream() for _ in range(max(1, int(streams)))]

    def step(self, n: int = 1):
        """n RK4 steps of every member: forked from the current stream onto the side streams and
        joined back, so work queued before and after stays ordered (and a graph capture records
        the members as parallel branches)."""
        cur = torch.cuda.current_stream()
        for s in self.streams:
            s.wait_stream(cur)
        for k, (sol, y, dt) in enumerate(zip(self.solvers, self.states, self.dts)):
            with torch.cuda.stream(self.streams[k % len(self.streams)]):
                sol.rk4(y, dt, n, self.precision)
        for s in self.streams:
            cur.wait_stream(s)

    def capture(self):
        """Capture one ensemble step (every member's RK4 step) in a HIP graph."""
        self.step(1)                      # scratch allocated and plans built before capture
        torch.cuda.synchronize()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self.step(1)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.graph = g
        return g

    def replay(self, n: int = 1):
        for _ in range(n):
            self.graph.replay()

    def close(self):
        self.graph = None
        for s in self.solvers:
            s.close()
        self.solvers = []
