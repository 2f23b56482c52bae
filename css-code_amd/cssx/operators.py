"""Directional compact derivatives on one subdomain, the counterpart of the Fortran module
``Derivatives`` (derivatives.f90): ``Initialize_Derivatives`` -> ``Operators(...)``,
``Dbydr/Dbydt/Dbydp`` -> ``dbydr/dbydt/dbydp``, ``D2bydr2/D2bydt2/D2bydp2`` ->
``d2bydr2/d2bydt2/d2bydp2``.

Fields are torch CUDA tensors of shape (nr, nphi, nth) in C order, i.e. the reference's
Fortran array f(nth, nphi, nr) with theta fastest.  Ghost buffers use the reference's halo
layouts (see include/cssx.h); at a physical wall slot 0 carries the prescribed first
derivative.  float64 tensors run the FP64 kernels, float32 tensors the FP32 kernels.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib


def _stream():
    return ctypes.c_void_p(int(torch.cuda.current_stream().cuda_stream))


class Operators:
    def __init__(self, nr, nth, nphi, dri, dti, dpi, r_walls=(True, True), th_walls=(True, True)):
        if not torch.cuda.is_available():
            raise RuntimeError("cssx.operators needs a ROCm GPU")
        self.shape = (nr, nphi, nth)
        g = _lib.Grid(nr, nth, nphi, int(r_walls[0]), int(r_walls[1]), int(th_walls[0]), int(th_walls[1]),
                      float(dri), float(dti), float(dpi))
        self._grid = g
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().cssx_ctx_create(ctypes.byref(g), ctypes.byref(h)), "cssx_ctx_create")
        self._ctx = h

    def close(self):
        if self._ctx is not None and self._ctx.value:
            _lib.lib().cssx_ctx_destroy(self._ctx)
        self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def deriv(self, f, axis, order=1, bdry=0, ghost_lo=None, ghost_hi=None, out=None):
        if tuple(f.shape) != self.shape or not f.is_cuda or not f.is_contiguous():
            raise ValueError(f"field must be a contiguous CUDA tensor of shape {self.shape}")
        if f.dtype == torch.float64:
            prec = _lib.PREC_F64
        elif f.dtype == torch.float32:
            prec = _lib.PREC_F32
        else:
            raise TypeError("field dtype must be float32 or float64")
        if out is None:
            out = torch.empty_like(f)
        for gbuf in (ghost_lo, ghost_hi):
            if gbuf is not None and (gbuf.dtype != f.dtype or not gbuf.is_cuda or not gbuf.is_contiguous()):
                raise ValueError("ghost buffers must be contiguous CUDA tensors of the field's dtype")
        rc = _lib.lib().cssx_deriv(self._ctx, axis, order, bdry, prec, f.data_ptr(), out.data_ptr(),
                                   ghost_lo.data_ptr() if ghost_lo is not None else None,
                                   ghost_hi.data_ptr() if ghost_hi is not None else None, _stream())
        _lib.check(rc, "cssx_deriv")
        return out

    def dbydr(self, f, bdry, ghost_lo=None, ghost_hi=None, out=None):
        return self.deriv(f, _lib.AXIS_R, 1, bdry, ghost_lo, ghost_hi, out)

    def dbydt(self, f, bdry, ghost_lo=None, ghost_hi=None, out=None):
        return self.deriv(f, _lib.AXIS_THETA, 1, bdry, ghost_lo, ghost_hi, out)

    def dbydp(self, f, ghost_lo=None, ghost_hi=None, out=None):
        return self.deriv(f, _lib.AXIS_PHI, 1, 0, ghost_lo, ghost_hi, out)

    def d2bydr2(self, f, bdry, ghost_lo=None, ghost_hi=None, out=None):
        return self.deriv(f, _lib.AXIS_R, 2, bdry, ghost_lo, ghost_hi, out)

    def d2bydt2(self, f, bdry, ghost_lo=None, ghost_hi=None, out=None):
        return self.deriv(f, _lib.AXIS_THETA, 2, bdry, ghost_lo, ghost_hi, out)

    def d2bydp2(self, f, ghost_lo=None, ghost_hi=None, out=None):
        return self.deriv(f, _lib.AXIS_PHI, 2, 0, ghost_lo, ghost_hi, out)
