"""GPU parity of the compact derivative kernels against the CPU oracle.

Every case calls the HIP kernels through the C ABI (cssx_compact_fd6 / cssx_deriv) on seeded
inputs and compares with the FP64 oracle restatement of derivatives.f90 / derivatives.py.
Tolerances: FP64 1e-12 of the result's max norm (the segmented solve reorders the Thomas
arithmetic; rounding only), FP32 2e-5.
"""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu

N_LIST = [8, 9, 16, 17, 61, 64, 100, 128, 257, 512, 1000, 1024]


@pytest.fixture(scope="module")
def cssx_mod():
    import cssx.derivatives as D
    import cssx.operators as O
    assert torch.cuda.is_available()
    return D, O


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(1e-300, np.abs(np.asarray(b)).max())


@pytest.mark.parametrize("n", N_LIST)
@pytest.mark.parametrize("dtype", [0, 1, 2, 3])
def test_compact_fd6_matches_python_restatement(cssx_mod, n, dtype):
    D, _ = cssx_mod
    rng = np.random.default_rng(1000 + n + dtype)
    for ibc in range(6):
        arr = rng.standard_normal(n)
        b1 = rng.standard_normal(4)
        b2 = rng.standard_normal(4)
        darr = np.zeros(n)
        darr[0], darr[-1] = rng.standard_normal(2)
        dri = 1.0 / 0.25
        ref, d_ref = oracle.compact_fd6_py(dri, arr, b1, b2, ibc, dtype, darr.copy())
        d_gpu = darr.copy()
        got = D.compact_fd6(dri, arr, b1, b2, ibc, dtype, d_gpu)
        assert _rel(got, ref) < 1e-12, (n, ibc, dtype, _rel(got, ref))
        assert _rel(d_gpu, d_ref) < 1e-12, (n, ibc, dtype)
        got2 = D.compact_fd6(dri, arr, b1, b2, ibc, dtype)      # darr omitted
        ref2, _ = oracle.compact_fd6_py(dri, arr, b1, b2, ibc, dtype, None)
        assert _rel(got2, ref2) < 1e-12


def test_compact_fd6_batched_axis(cssx_mod):
    D, _ = cssx_mod
    rng = np.random.default_rng(7)
    n = 96
    arr = rng.standard_normal((5, n, 3))
    got = D.compact_fd6(4.0, arr, 0.0, 0.0, 1, 0, axis=1)
    for i in range(5):
        for j in range(3):
            ref, _ = oracle.compact_fd6_py(4.0, arr[i, :, j], 0.0, 0.0, 1, 0)
            assert _rel(got[i, :, j], ref) < 1e-12
    t = torch.as_tensor(arr, device="cuda")
    got_t = D.compact_fd6(4.0, t, 0.0, 0.0, 1, 0, axis=1)
    assert isinstance(got_t, torch.Tensor) and got_t.is_cuda
    assert _rel(got_t.cpu().numpy(), got) < 1e-15


def test_compact_fd6_reference_known_answer(cssx_mod):
    """CSS-Python/Unit-Tests/test_compact_diff.py:18-67 with nr=128, ibc=1, dtype=0."""
    D, _ = cssx_mod
    nr, r1, r2 = 128, 0.75, 1.0
    dr = (r2 - r1) / (nr - 1)
    r = dr * np.arange(nr) + r1
    f = np.sin(2 * np.pi * (r - r1) / (r2 - r1))
    exact = 2 * np.pi * np.cos(2 * np.pi * (r - r1) / (r2 - r1)) / (r2 - r1)
    got = D.compact_fd6(1 / (r2 - r1), f, 0.0, 0.0, 1, 0)
    assert np.sqrt(dr * np.sum((got - exact) ** 2)) < 1e-6


def _ghost_buffers(axis, shape, rng, lo_wall, hi_wall):
    nr, nphi, nth = shape
    if axis == 0:
        gshape = (4, nphi, nth)
    elif axis == 1:
        gshape = (nr, nphi, 4)
    else:
        gshape = (nr, 4, nth)
    glo = rng.standard_normal(gshape)
    ghi = rng.standard_normal(gshape)
    return glo, ghi


CASES = []
for _axis in (0, 1, 2):
    for _order in (1, 2):
        for _lo, _hi in ([(1, 1), (0, 0), (1, 0), (0, 1)] if _axis < 2 else [(0, 0), (None, None)]):
            for _bdry in ((0, 1, 2, 3, 4) if _axis < 2 and (_lo or _hi) else (0,)):
                CASES.append((_axis, _order, _lo, _hi, _bdry))


@pytest.mark.parametrize("shape", [(24, 40, 72), (40, 16, 130)])
@pytest.mark.parametrize("case", CASES)
def test_deriv_field_matches_oracle(cssx_mod, shape, case):
    _, O = cssx_mod
    axis, order, lo, hi, bdry = case
    nr, nphi, nth = shape
    rng = np.random.default_rng(hash((shape, case)) % (2 ** 32))
    f = rng.standard_normal(shape)
    dri, dti, dpi = 1.3 * (nr - 1), 0.7 * (nth - 1), 2.1 * nphi
    periodic = axis == 2 and lo is None
    if axis == 0:
        walls_r, walls_t = (bool(lo), bool(hi)), (True, True)
    elif axis == 1:
        walls_r, walls_t = (True, True), (bool(lo), bool(hi))
    else:
        walls_r, walls_t = (True, True), (True, True)
    ops = O.Operators(nr, nth, nphi, dri, dti, dpi, r_walls=walls_r, th_walls=walls_t)
    glo, ghi = _ghost_buffers(axis, shape, rng, lo, hi)
    if periodic:
        glo = ghi = None
    scale = [dri, dti, dpi][axis]
    sc = scale if order == 1 else scale * scale
    hneu = scale if order == 1 else dri
    ref = oracle.deriv_field(axis, order, f, lo_wall=bool(lo), hi_wall=bool(hi), bdry=bdry, ghost_lo=glo,
                             ghost_hi=ghi, scale=sc, hneu=hneu)
    for prec, tol in ((torch.float64, 1e-12), (torch.float32, 3e-5)):
        ft = torch.as_tensor(f, dtype=prec, device="cuda")
        gl = None if glo is None else torch.as_tensor(glo, dtype=prec, device="cuda")
        gh = None if ghi is None else torch.as_tensor(ghi, dtype=prec, device="cuda")
        got = ops.deriv(ft, axis, order, bdry, gl, gh).double().cpu().numpy()
        err = _rel(got, ref)
        assert err < tol, (case, prec, err)
    ops.close()


def test_deriv_c1_recipe(cssx_mod):
    """The plain-derivative workload (64 x 64 x 128, tanh-stretched r on [0.72, 0.97], the
    seeded 8-term field of tests/c1_field.py, phi halos taken from the field): cssx_deriv on
    every axis against the FP64 oracle (128 eps max|f| scale) and the analytic derivatives (the
    bounds the oracle itself meets), with one-sided walls (code 1) and with Neumann walls
    carrying the exact wall derivative (code 4); and the compact_fd6 drop-in on every r and
    theta pencil and on sampled phi pencils (dtype 3, halo ghosts) against the CSS-Python
    restatement."""
    import c1_field
    D, O = cssx_mod
    c = c1_field.build()
    f = c["f"]
    nr, nphi, nth = f.shape
    dxdr = c["dxdr"][:, None, None]
    ref = {
        0: oracle.deriv_field(0, 1, f, bdry=1, scale=c["dri"]),
        1: oracle.deriv_field(1, 1, f, bdry=1, scale=c["dti"]),
        2: oracle.deriv_field(2, 1, f, lo_wall=False, hi_wall=False, ghost_lo=c["ghost_lo"],
                              ghost_hi=c["ghost_hi"], scale=c["dpi"]),
    }
    exact = {0: c["fr"], 1: c["ft"], 2: c["fp"]}
    bound = {0: 5e-6, 1: 1e-5, 2: 1e-9}
    ops = O.Operators(nr, nth, nphi, c["dri"], c["dti"], c["dpi"])
    # Rounding bound: 128 eps(T) max|f| scale, absolute.  The Dirichlet-type wall rows combine
    # values with coefficient sums near 46 (2 (10/3 + 3 + 6 + 1/3) + 8 (43/96 + 5/6 + 9/8 + 1/6
    # + 1/96)) and the solve carries that rounding on at most ~3x; for this smooth field the
    # index-space r derivative is only ~0.02 |f|, so the rounding is large relative to the
    # result (measured on the r axis: 72 eps max|f| scale in FP64, 46 in FP32, of which the FP64
    # oracle's own response to FP32-rounded input is 2).
    scales = {0: c["dri"], 1: c["dti"], 2: c["dpi"]}
    fmax = np.abs(f).max()
    for prec in (torch.float64, torch.float32):
        eps = float(torch.finfo(prec).eps)
        x = torch.as_tensor(f, dtype=prec, device="cuda")
        gl = torch.as_tensor(c["ghost_lo"], dtype=prec, device="cuda")
        gh = torch.as_tensor(c["ghost_hi"], dtype=prec, device="cuda")
        got = {0: ops.dbydr(x, 1), 1: ops.dbydt(x, 1), 2: ops.dbydp(x, gl, gh)}
        for ax in range(3):
            g = got[ax].double().cpu().numpy()
            lim = 128 * eps * fmax * scales[ax]
            err = np.abs(g - ref[ax]).max()
            assert err < lim, (ax, prec, err, lim)
            tol = lim / np.abs(ref[ax]).max()
            phys = g * dxdr if ax == 0 else g
            assert _rel(phys, exact[ax]) < bound[ax] + 2 * tol, (ax, prec, _rel(phys, exact[ax]))
    # Neumann walls with the exact wall derivatives in slot 0 of the r ghost buffers
    gl = np.zeros((4, nphi, nth))
    gh = np.zeros((4, nphi, nth))
    gl[0] = c["fr"][0] / c["dxdr"][0]
    gh[0] = c["fr"][-1] / c["dxdr"][-1]
    r4 = oracle.deriv_field(0, 1, f, bdry=4, ghost_lo=gl, ghost_hi=gh, scale=c["dri"], hneu=c["dri"])
    g4 = ops.dbydr(torch.as_tensor(f, device="cuda"), 4, torch.as_tensor(gl, device="cuda"),
                   torch.as_tensor(gh, device="cuda")).cpu().numpy()
    eps64 = float(np.finfo(np.float64).eps)
    assert np.abs(g4 - r4).max() < 128 * eps64 * fmax * c["dri"]
    assert _rel(g4 * dxdr, c["fr"]) < 2e-8
    ops.close()
    # the CSS-Python drop-in: r pencils (axis 0) and theta pencils (axis 2), walls, ibc 1
    pr_ = D.compact_fd6(1.0, f, 0.0, 0.0, 1, 0, axis=0)
    assert np.abs(pr_ - ref[0]).max() < 128 * eps64 * fmax * c["dri"]
    pt_ = D.compact_fd6(1.0 / (np.pi / 2), f, 0.0, 0.0, 1, 0, axis=2)
    assert np.abs(pt_ - ref[1]).max() < 128 * eps64 * fmax * c["dti"]
    # phi pencils with their halos (dtype 3: the Python's halo-row scaling included)
    pp_ = D.compact_fd6(1.0 / (np.pi / 2), f, c["ghost_lo"], c["ghost_hi"], 0, 3, axis=1)
    rng = np.random.default_rng(12)
    for _ in range(24):
        ir, it = rng.integers(0, nr), rng.integers(0, nth)
        want, _ = oracle.compact_fd6_py(1.0 / (np.pi / 2), f[ir, :, it], c["ghost_lo"][ir, :, it],
                                       c["ghost_hi"][ir, :, it], 0, 3)
        assert np.abs(pp_[ir, :, it] - want).max() < 128 * eps64 * fmax * c["dpi"]


LONG_CASES = []
for _axis, _shapes in ((0, [(1024, 8, 16), (600, 8, 12)]), (1, [(8, 8, 1024), (8, 9, 600)]),
                       (2, [(8, 1024, 16), (8, 600, 12), (8, 513, 16)])):
    for _shape in _shapes:
        for _order in (1, 2):
            for _lo, _hi in ([(1, 1), (0, 0), (1, 0)] if _axis < 2 else [(0, 0), (None, None)]):
                LONG_CASES.append((_shape, (_axis, _order, _lo, _hi, 1 if _axis < 2 and (_lo or _hi) else 0)))


@pytest.mark.parametrize("shape,case", LONG_CASES)
def test_deriv_long_pencils_match_oracle(cssx_mod, shape, case):
    """Pencils longer than 512 rows on every axis (the 256x512x1024 workload's phi axis is 1024
    long, periodic): walls, halo ghosts and the periodic wrap against the oracle."""
    _, O = cssx_mod
    axis, order, lo, hi, bdry = case
    print(f"long pencil {shape} {case}", flush=True)
    nr, nphi, nth = shape
    rng = np.random.default_rng(hash((shape, case, 7)) % (2 ** 32))
    f = rng.standard_normal(shape)
    dri, dti, dpi = 1.3 * (nr - 1), 0.7 * (nth - 1), 2.1 * nphi
    periodic = axis == 2 and lo is None
    walls_r = (bool(lo), bool(hi)) if axis == 0 else (True, True)
    walls_t = (bool(lo), bool(hi)) if axis == 1 else (True, True)
    ops = O.Operators(nr, nth, nphi, dri, dti, dpi, r_walls=walls_r, th_walls=walls_t)
    glo, ghi = (None, None) if periodic else _ghost_buffers(axis, shape, rng, lo, hi)
    scale = [dri, dti, dpi][axis]
    sc = scale if order == 1 else scale * scale
    hneu = scale if order == 1 else dri
    ref = oracle.deriv_field(axis, order, f, lo_wall=bool(lo), hi_wall=bool(hi), bdry=bdry, ghost_lo=glo,
                             ghost_hi=ghi, scale=sc, hneu=hneu)
    for prec, tol in ((torch.float64, 1e-12), (torch.float32, 3e-5)):
        ft = torch.as_tensor(f, dtype=prec, device="cuda")
        gl = None if glo is None else torch.as_tensor(glo, dtype=prec, device="cuda")
        gh = None if ghi is None else torch.as_tensor(ghi, dtype=prec, device="cuda")
        got = ops.deriv(ft, axis, order, bdry, gl, gh).double().cpu().numpy()
        err = _rel(got, ref)
        assert err < tol, (shape, case, prec, err)
    ops.close()
