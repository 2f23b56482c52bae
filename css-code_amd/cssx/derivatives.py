"""Drop-in for CSS-Python's ``derivatives`` module, computed by the HIP kernels.

``compact_fd6(dri, arr, b1, b2, ibc, dtype=0, darr=None)`` keeps the signature and the
semantics of CSS-Python/derivatives.py:32-185:

* ``dri = 1/(x2-x1)``; the rows are built in units of ``dli = n-1`` (``n`` if ``ibc == 5``)
  and the result is ``solution * dri``;
* ``b1``/``b2`` are ignored scalars for walls or 4 ghost values (``b1[3]`` and ``b2[0]``
  adjacent to the data) for ``dtype`` 1/3 and 2/3;
* Neumann values ``f'/dri`` are preloaded in ``darr[0]`` / ``darr[-1]``; ``darr`` is
  mutated in place and holds the assembled right-hand side afterwards;
* the halo rows of ``dtype`` 1/2/3 carry the Python's extra ``dli/60`` factor on their
  stencil part (exactly the Fortran rows when ``n == 61``).

Extensions over the reference: ``arr`` may be N-dimensional (``axis`` selects the pencil
direction, every other index is an independent pencil solved in the same launch), a
``torch`` CUDA tensor (no host round trip), and ``darr is None`` is tested with ``is``.
Inputs that live on the host are copied to the GPU and the result copied back; the
arithmetic always runs in libcssx.so.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib


def _device():
    if not torch.cuda.is_available():
        raise RuntimeError("cssx.derivatives needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda")


def _to_dev(x, dev):
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=torch.float64)
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=dev)


def _stream_ptr():
    return ctypes_void(torch.cuda.current_stream().cuda_stream)


def ctypes_void(v):
    import ctypes
    return ctypes.c_void_p(int(v))


def compact_fd6(dri, arr, b1, b2, ibc, dtype=0, darr=None, axis=-1):
    """Compact 6th-order first derivative (CSS-Python/derivatives.py:32-185) on the GPU."""
    dev = _device()
    is_torch = isinstance(arr, torch.Tensor)
    a = _to_dev(arr, dev)
    if a.ndim == 0:
        raise ValueError("compact_fd6: arr must have at least one dimension")
    ax = axis % a.ndim
    a_m = torch.movedim(a, ax, 0)
    n = a_m.shape[0]
    rest = tuple(a_m.shape[1:])
    npen = int(np.prod(rest)) if rest else 1
    a2 = a_m.reshape(n, npen).contiguous()
    if n < 5:
        raise ValueError(f"compact_fd6: pencil length {n} < 5 (the rows reach arr[4] and arr[n-5])")
    dtype = int(dtype)
    ibc = int(ibc)

    def ghosts(b, needed):
        if not needed:
            return None
        bt = _to_dev(b, dev)
        if bt.numel() == 4:
            return bt.reshape(4, 1).expand(4, npen).contiguous()
        # per-pencil ghosts: shaped like arr with 4 along `axis`, or (4,) + the pencil shape
        gax = ax if bt.ndim == a.ndim else 0
        if bt.ndim == 0 or bt.shape[gax] != 4 or bt.numel() != 4 * npen:
            raise ValueError("compact_fd6: dtype %d needs 4 ghost values per pencil" % dtype)
        return torch.movedim(bt, gax, 0).reshape(4, npen).contiguous()

    g1 = ghosts(b1, dtype in (1, 3))
    g2 = ghosts(b2, dtype in (2, 3))
    if darr is not None:
        d_in = _to_dev(darr, dev)
        d2 = torch.movedim(d_in, ax, 0).reshape(n, npen).contiguous().clone()
    else:
        d2 = torch.zeros((n, npen), dtype=torch.float64, device=dev)
    res = torch.empty((n, npen), dtype=torch.float64, device=dev)
    L = _lib.lib()
    rc = L.cssx_compact_fd6(n, npen, float(dri), ibc, dtype, a2.data_ptr(),
                            g1.data_ptr() if g1 is not None else None,
                            g2.data_ptr() if g2 is not None else None,
                            d2.data_ptr(), res.data_ptr(), _stream_ptr())
    _lib.check(rc, "cssx_compact_fd6")
    out = torch.movedim(res.reshape((n,) + rest), 0, ax)
    if darr is not None:
        d_after = torch.movedim(d2.reshape((n,) + rest), 0, ax)
        if isinstance(darr, torch.Tensor):
            darr.copy_(d_after.to(darr.dtype))
        else:
            np.copyto(darr, d_after.cpu().numpy().astype(np.asarray(darr).dtype, copy=False))
    if is_torch:
        return out
    return out.cpu().numpy()
