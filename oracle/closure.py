"""CPU oracle of the learned-closure kernels (test infrastructure only: imported by tests/ and
nothing else; the product path is css-code_amd/csrc/cssx_closure.hip).

The closure has no counterpart in the reference (it stands beside diffusion.f90's SLD as a
coarse-grid subgrid model), so nothing here restates a reference routine; each function
restates the mathematics the HIP kernels implement, in numpy FP64:

  bf16          round-to-nearest-even to bfloat16 (the matrix cores' operand format)
  gelu, dgelu   0.5 z (1 + erf(z/sqrt 2)) and its derivative Phi(z) + z phi(z)
  forward       216 -> h1 -> h2 -> 8 MLP with the GPU's BF16 rounding points (inputs, weights,
                hidden activations), FP64 sums
  grads         backward of the mean squared error with the same rounding points (dY, dZ2, dZ1
                rounded to BF16 where they feed a matrix product; bias gradients from the
                unrounded values); exact=True drops every rounding (pinned by finite
                differences in tests/test_oracle_closure.py)
  adam          Adam (beta 0.9 / 0.999, eps 1e-8, bias-corrected)
  box_filter    mean over f^3 blocks
  features      3x3x3 stencils of the 8 fields (walls clamp, phi wraps), rho/s as fluctuations

Parity status: every function is pinned by a closed form, an invariant or finite differences
in tests/test_oracle_closure.py.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

NFEAT = 216
NOUT = 8


def bf16(x):
    """Round float32 values to the nearest bfloat16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).reshape(a.shape)


def gelu(z):
    z = np.asarray(z, dtype=np.float64)
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def dgelu(z):
    z = np.asarray(z, dtype=np.float64)
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def unpack(params, h1, h2, nin=NFEAT, nout=NOUT):
    """Flat parameter vector -> W1 (nin,h1), b1, W2 (h1,h2), b2, W3 (h2,nout), b3."""
    p = np.asarray(params, dtype=np.float64)
    o = 0
    out = []
    for shape in ((nin, h1), (h1,), (h1, h2), (h2,), (h2, nout), (nout,)):
        k = int(np.prod(shape))
        out.append(p[o:o + k].reshape(shape))
        o += k
    assert o == p.size
    return out


def pack(parts):
    return np.concatenate([np.asarray(q, dtype=np.float64).ravel() for q in parts])


def forward(params, X, h1, h2, exact=False):
    """Returns Y and the intermediates (Z1, H1, Z2, H2).  X rows are the (BF16) inputs."""
    W1, b1, W2, b2, W3, b3 = unpack(params, h1, h2)
    r = (lambda a: np.asarray(a, np.float64)) if exact else (lambda a: bf16(a).astype(np.float64))
    Xb = r(X)
    Z1 = Xb @ r(W1) + b1
    H1 = r(gelu(Z1))
    Z2 = H1 @ r(W2) + b2
    H2 = r(gelu(Z2))
    Y = H2 @ r(W3) + b3
    return Y, (Xb, Z1, H1, Z2, H2)


def loss(params, X, T, h1, h2, exact=False):
    Y, _ = forward(params, X, h1, h2, exact)
    return float(np.mean((Y - np.asarray(T, np.float64)) ** 2))


def grads(params, X, T, h1, h2, exact=False):
    """(loss, flat gradient) of the mean squared error over rows x 8 outputs."""
    W1, b1, W2, b2, W3, b3 = unpack(params, h1, h2)
    r = (lambda a: np.asarray(a, np.float64)) if exact else (lambda a: bf16(a).astype(np.float64))
    Y, (Xb, Z1, H1, Z2, H2) = forward(params, X, h1, h2, exact)
    T = np.asarray(T, np.float64)
    n = Y.size
    dY = 2.0 * (Y - T) / n
    if not exact:
        dY = dY.astype(np.float32).astype(np.float64)
    dYb = r(dY)
    dW3 = H2.T @ dYb
    db3 = dY.sum(0)
    dZ2 = (dYb @ r(W3).T) * dgelu(Z2)
    if not exact:
        dZ2 = dZ2.astype(np.float32).astype(np.float64)
    dZ2b = r(dZ2)
    dW2 = H1.T @ dZ2b
    db2 = dZ2.sum(0)
    dZ1 = (dZ2b @ r(W2).T) * dgelu(Z1)
    if not exact:
        dZ1 = dZ1.astype(np.float32).astype(np.float64)
    dZ1b = r(dZ1)
    dW1 = Xb.T @ dZ1b
    db1 = dZ1.sum(0)
    return float(np.mean((Y - T) ** 2)), pack([dW1, db1, dW2, db2, dW3, db3])


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (t counts from 1); returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def box_filter(state, f):
    """(nv, nr, nphi, nth) -> (nv, nr/f, nphi/f, nth/f), mean over f^3 blocks."""
    s = np.asarray(state, dtype=np.float64)
    nv, nr, nphi, nth = s.shape
    return s.reshape(nv, nr // f, f, nphi // f, f, nth // f, f).mean(axis=(2, 4, 6))


def features(state, rbar=None, sbar=None, fluct=False):
    """Raw (unstandardized) features (N, 216) of an 8-field state (8, nr, nphi, nth)."""
    s = np.array(state, dtype=np.float64)
    nv, nr, nphi, nth = s.shape
    assert nv == 8
    if not fluct:
        s[0] -= np.asarray(rbar, np.float64)[:, None, None]
        s[4] -= np.asarray(sbar, np.float64)[:, None, None]
    ir, ip, it = np.meshgrid(np.arange(nr), np.arange(nphi), np.arange(nth), indexing="ij")
    cols = []
    for iv in range(8):
        for dr in (-1, 0, 1):
            for dp in (-1, 0, 1):
                for dt in (-1, 0, 1):
                    jr = np.clip(ir + dr, 0, nr - 1)
                    jp = (ip + dp) % nphi
                    jt = np.clip(it + dt, 0, nth - 1)
                    cols.append(s[iv][jr, jp, jt].ravel())
    return np.stack(cols, axis=1)
