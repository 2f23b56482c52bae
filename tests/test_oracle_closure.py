"""Pins of the learned-closure oracle (oracle/closure.py) on the CPU.

Each function is checked against a library routine or a closed form that a plausible slip
(a dropped term, a transposed operand, a wrong rounding rule) would fail:
  * bf16 rounding: hand-worked ties and torch's float32 -> bfloat16 conversion;
  * GELU: torch's exact (erf) GELU, the identity gelu(x) - gelu(-x) = x, and finite differences
    of the derivative;
  * backward: torch autograd on the same FP64 network, and central finite differences;
  * Adam: torch.optim.Adam over several steps, and the closed form of the first step (|dp| = lr);
  * box filter: constants and index-linear fields; features: a brute-force loop on a tiny grid.
"""
import math

import numpy as np
import pytest
import torch

from oracle import closure as oc

H1, H2 = 16, 12


def _params(seed, h1=H1, h2=H2, scale=0.3):
    rng = np.random.default_rng(seed)
    n = oc.NFEAT * h1 + h1 + h1 * h2 + h2 + h2 * oc.NOUT + oc.NOUT
    return (rng.standard_normal(n) * scale).astype(np.float32).astype(np.float64)


def test_bf16_rounding_ties_and_torch():
    one = np.float32(1.0)
    assert oc.bf16(np.float32(1.0)) == one
    assert oc.bf16(np.float32(1.0 + 2.0 ** -8)) == one                        # tie -> even (down)
    assert oc.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == np.float32(1.0 + 2.0 ** -6)   # tie -> even (up)
    assert oc.bf16(np.float32(1.0 + 2.0 ** -8 + 2.0 ** -20)) == np.float32(1.0 + 2.0 ** -7)
    x = np.random.default_rng(0).standard_normal(100000).astype(np.float32) * 10
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(oc.bf16(x), ref)


def test_gelu_matches_torch_and_identities():
    z = np.linspace(-8, 8, 2001)
    ref = torch.nn.functional.gelu(torch.from_numpy(z)).numpy()
    assert np.abs(oc.gelu(z) - ref).max() < 1e-14
    assert np.abs(oc.gelu(z) - oc.gelu(-z) - z).max() < 1e-13
    h = 1e-6
    fd = (oc.gelu(z + h) - oc.gelu(z - h)) / (2 * h)
    assert np.abs(oc.dgelu(z) - fd).max() < 1e-8


def test_exact_gradient_matches_autograd_and_finite_differences():
    rng = np.random.default_rng(3)
    p = _params(1)
    X = rng.standard_normal((7, oc.NFEAT))
    T = rng.standard_normal((7, oc.NOUT))
    L, g = oc.grads(p, X, T, H1, H2, exact=True)
    # autograd on the same network
    pt = torch.tensor(p, requires_grad=True)
    W1, b1, W2, b2, W3, b3 = oc.unpack(np.zeros_like(p), H1, H2)
    o = 0
    parts = []
    for a in (W1, b1, W2, b2, W3, b3):
        parts.append(pt[o:o + a.size].reshape(a.shape))
        o += a.size
    gl = torch.nn.functional.gelu
    Xt = torch.from_numpy(X)
    Y = gl(gl(Xt @ parts[0] + parts[1]) @ parts[2] + parts[3]) @ parts[4] + parts[5]
    Lt = ((Y - torch.from_numpy(T)) ** 2).mean()
    Lt.backward()
    assert abs(L - Lt.item()) < 1e-12 * max(1.0, abs(L))
    assert np.abs(g - pt.grad.numpy()).max() < 1e-11 * np.abs(g).max()
    # central differences on a sample of parameters
    idx = rng.choice(p.size, 40, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (oc.loss(p + e, X, T, H1, H2, exact=True) - oc.loss(p - e, X, T, H1, H2, exact=True)) / (2 * h)
        assert abs(fd - g[i]) < 1e-7 + 1e-5 * abs(g[i])


def test_rounded_gradient_close_to_exact():
    rng = np.random.default_rng(4)
    p = _params(2)
    X = oc.bf16(rng.standard_normal((64, oc.NFEAT)).astype(np.float32))
    T = rng.standard_normal((64, oc.NOUT))
    _, ge = oc.grads(p, X, T, H1, H2, exact=True)
    _, gr = oc.grads(p, X, T, H1, H2, exact=False)
    assert np.abs(gr - ge).max() < 3e-2 * np.abs(ge).max()
    assert np.abs(gr - ge).max() > 0      # the rounding points are really there


def test_adam_matches_torch_and_first_step():
    rng = np.random.default_rng(5)
    p0 = rng.standard_normal(50)
    gs = [rng.standard_normal(50) for _ in range(5)]
    pt = torch.tensor(p0, requires_grad=True)
    opt = torch.optim.Adam([pt], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for t, g in enumerate(gs, 1):
        opt.zero_grad()
        pt.grad = torch.tensor(g)
        opt.step()
        p, m, v = oc.adam(p, g, m, v, t, 1e-2)
        if t == 1:   # bias-corrected first step moves every parameter by lr against its gradient
            assert np.allclose(p - p0, -1e-2 * np.sign(g), rtol=1e-4, atol=0)
    assert np.abs(p - pt.detach().numpy()).max() < 1e-12


def test_box_filter_constants_and_linear_fields():
    nv, nr, nphi, nth, f = 2, 8, 12, 16, 4
    c = np.full((nv, nr, nphi, nth), 3.25)
    assert np.allclose(oc.box_filter(c, f), 3.25)
    ir, ip, it = np.meshgrid(np.arange(nr), np.arange(nphi), np.arange(nth), indexing="ij")
    lin = np.stack([1.0 * ir + 10 * ip + 100 * it, 2.0 * it])
    b = oc.box_filter(lin, f)
    cr, cp, ct = np.meshgrid(np.arange(nr // f), np.arange(nphi // f), np.arange(nth // f), indexing="ij")
    ctr = lambda k: f * k + (f - 1) / 2.0   # mean index of a block
    assert np.allclose(b[0], ctr(cr) + 10 * ctr(cp) + 100 * ctr(ct))
    assert np.allclose(b[1], 2.0 * ctr(ct))
    assert np.isclose(b.mean(), lin.mean())


def test_features_brute_force():
    rng = np.random.default_rng(6)
    nr, nphi, nth = 3, 4, 5
    st = rng.standard_normal((8, nr, nphi, nth))
    rbar = rng.standard_normal(nr)
    sbar = rng.standard_normal(nr)
    F = oc.features(st, rbar, sbar)
    assert F.shape == (nr * nphi * nth, 216)
    for pt in (0, 7, 31, nr * nphi * nth - 1):
        ir, rest = divmod(pt, nphi * nth)
        ip, it = divmod(rest, nth)
        q = 0
        for iv in range(8):
            for dr in (-1, 0, 1):
                for dp in (-1, 0, 1):
                    for dt in (-1, 0, 1):
                        jr = min(max(ir + dr, 0), nr - 1)
                        jp = (ip + dp + nphi) % nphi
                        jt = min(max(it + dt, 0), nth - 1)
                        v = st[iv, jr, jp, jt] - (rbar[jr] if iv == 0 else sbar[jr] if iv == 4 else 0.0)
                        assert F[pt, q] == v
                        q += 1
    fl = st.copy()
    fl[0] -= rbar[:, None, None]
    fl[4] -= sbar[:, None, None]
    assert np.array_equal(oc.features(fl, fluct=True), F)
