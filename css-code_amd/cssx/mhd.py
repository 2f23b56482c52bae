"""MHD right-hand side and RK4 stepping on the GPU: the counterpart of Module Physics
(rhs_mpi.f90: RHS, Bound, Compute_Step_Size) and Module Time_Step_RK4 (step_rk4.f90:
Step_RK4) for one subdomain.

States are torch CUDA tensors of shape (nv, nr, nphi, nth) in C order (the reference's
vars(nth, nphi, nr, nv)), variables rho, w=u_theta, v=u_phi, u=u_r, s, A_theta, A_phi, A_r.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .model import Config, Model


def _stream():
    return ctypes.c_void_p(int(torch.cuda.current_stream().cuda_stream))


class Solver:
    """GPU solver for one subdomain (walls at both r and theta ends, periodic phi)."""

    def __init__(self, model: Model | Config):
        if isinstance(model, Config):
            model = Model(model)
        if not torch.cuda.is_available():
            raise RuntimeError("cssx.mhd needs a ROCm GPU")
        self.model = model
        c = model.c
        self.shape = (c.nv, c.nr, c.nphi, c.nth)
        self._g = model.grid_struct()
        self._p = model.params_struct()
        self._pr = model.profiles_struct()
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().cssx_mhd_create(ctypes.addressof(self._g), ctypes.addressof(self._p),
                                              ctypes.addressof(self._pr), ctypes.byref(h)), "cssx_mhd_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().cssx_mhd_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check_state(self, x, dtype=None):
        if tuple(x.shape) != self.shape or not x.is_cuda or not x.is_contiguous():
            raise ValueError(f"state must be a contiguous CUDA tensor of shape {self.shape}")
        if dtype is not None and x.dtype != dtype:
            raise TypeError(f"state dtype must be {dtype}")

    @staticmethod
    def _prec(x):
        if x.dtype == torch.float64:
            return _lib.PREC_F64
        if x.dtype == torch.float32:
            return _lib.PREC_F32
        raise TypeError("float32 or float64 state expected")

    def rhs(self, state, out=None, want_b=False):
        """RHS(call_num): d(state)/dt; with want_b also B = curl A (Compute_B).
        float32 states are in fluctuation form (see to_f32)."""
        self._check_state(state)
        if out is None:
            out = torch.empty_like(state)
        b = torch.empty((3,) + self.shape[1:], dtype=state.dtype, device=state.device) if want_b else None
        rc = _lib.lib().cssx_rhs(self._h, self._prec(state), state.data_ptr(), out.data_ptr(),
                                 b.data_ptr() if b is not None else None, _stream())
        _lib.check(rc, "cssx_rhs")
        return (out, b) if want_b else out

    def background(self, device="cuda"):
        """(nv, nr, 1, 1) FP64 tensor with rho_bar and s_bar in slots 0 and 4, zeros elsewhere."""
        m = self.model
        bg = torch.zeros((self.shape[0], self.shape[1], 1, 1), dtype=torch.float64, device=device)
        bg[0, :, 0, 0] = torch.as_tensor(m.rho_bar, device=device)
        bg[4, :, 0, 0] = torch.as_tensor(m.s_bar, device=device)
        return bg

    def to_f32(self, y):
        """FP64 full state -> FP32 state in fluctuation form (rho - rho_bar, s - s_bar)."""
        return (y - self.background(y.device)).float().contiguous()

    def from_f32(self, y32):
        """FP32 fluctuation-form state -> FP64 full state."""
        return (y32.double() + self.background(y32.device)).contiguous()

    def diffusion(self, state, b=None, out=None):
        """Compute_Diffusion(1): the SLD tendency of a state (A in slots 6-8; float32 states in
        fluctuation form).  b = curl A as RHS(1) leaves it; computed when not given."""
        self._check_state(state)
        if out is None:
            out = torch.empty_like(state)
        if b is not None and (tuple(b.shape) != (3,) + self.shape[1:] or b.dtype != state.dtype
                              or not b.is_contiguous()):
            raise ValueError("b must be a contiguous (3, nr, nphi, nth) tensor of the state's dtype")
        rc = _lib.lib().cssx_diffusion(self._h, self._prec(state), state.data_ptr(),
                                       b.data_ptr() if b is not None else None, out.data_ptr(), _stream())
        _lib.check(rc, "cssx_diffusion")
        return out

    def bound(self, state):
        """Bound(): wall conditions applied in place."""
        self._check_state(state)
        _lib.check(_lib.lib().cssx_bound(self._h, self._prec(state), state.data_ptr(), _stream()), "cssx_bound")
        return state

    def rk4(self, y0, dt, nsteps=1, precision="f64", scheme="rk4"):
        """nsteps of Step_RK4 on the FP64 state y0 (in place).  precision='f32' runs the stages,
        RHS and diffusion in FP32 with FP64 accumulation.  scheme 'rk3_tvd' / 'rk2_tvd' runs
        Step_RK3_TVD / Step_RK2_TVD instead."""
        self._check_state(y0, torch.float64)
        if precision not in ("f64", "f32"):
            raise ValueError("precision must be 'f64' or 'f32'")
        prec = _lib.PREC_F64 if precision == "f64" else _lib.PREC_F32
        name = {"rk4": "cssx_rk4", "rk3_tvd": "cssx_rk3_tvd", "rk2_tvd": "cssx_rk2_tvd"}[scheme]
        _lib.check(getattr(_lib.lib(), name)(self._h, prec, y0.data_ptr(), float(dt), int(nsteps), _stream()), name)
        return y0

    def capture(self, y0, dt, nsteps=1, precision="f64", scheme="rk4"):
        """Capture `nsteps` steps of `scheme` on the FP64 state y0 (in place) in a HIP graph
        (torch.cuda.CUDAGraph holds it); ``StepGraph.replay(n)`` then advances y0 by n * nsteps
        steps with one graph launch each.  One eager step runs first, on a copy of y0, so that
        the scratch buffers and plans exist before capture (the capture itself allocates
        nothing); y0 is left unchanged until the first replay.  An attached closure runs on the
        captured stream (no side-stream fork under capture)."""
        self._check_state(y0, torch.float64)
        warm = y0.clone()
        self.rk4(warm, dt, 1, precision, scheme)
        torch.cuda.synchronize()
        del warm
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self.rk4(y0, dt, nsteps, precision, scheme)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        return StepGraph(g, y0, nsteps)

    def cfl(self, y0, cfl_safety=None):
        """Compute_Step_Size: the CFL time step (synchronizes)."""
        self._check_state(y0, torch.float64)
        dt = ctypes.c_double()
        cs = self.model.c.cfl_safety if cfl_safety is None else cfl_safety
        _lib.check(_lib.lib().cssx_cfl(self._h, y0.data_ptr(), float(cs), ctypes.byref(dt), _stream()), "cssx_cfl")
        return dt.value

    def conservation(self, y0):
        """{E, L, m, mrsint}: the totals Compute_Step_Size reports (rhs_mpi.f90:156-216) for an
        FP64 state (synchronizes)."""
        self._check_state(y0, torch.float64)
        out = (ctypes.c_double * 4)()
        _lib.check(_lib.lib().cssx_conservation(self._h, y0.data_ptr(), out, _stream()), "cssx_conservation")
        return {"E": out[0], "L": out[1], "m": out[2], "mrsint": out[3]}


class StepGraph:
    """A captured sequence of RK4 steps of one state (Solver.capture)."""

    def __init__(self, graph, state, nsteps):
        self.graph = graph
        self.state = state          # the captured kernels read and write this tensor
        self.nsteps = nsteps

    def replay(self, n: int = 1):
        """n launches of the graph on the current stream: n * nsteps steps of the state."""
        for _ in range(n):
            self.graph.replay()
        return self.state
