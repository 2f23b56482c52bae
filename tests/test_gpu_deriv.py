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


def test_deriv_analytic_sphere_c1(cssx_mod):
    """Config-1 style field on a uniform grid: derivatives against the analytic ones."""
    _, O = cssx_mod
    nr, nth, nphi = 64, 64, 128
    r = np.linspace(0.72, 0.97, nr)
    th = np.linspace(np.pi / 4, 3 * np.pi / 4, nth)
    ph = np.arange(nphi) * (np.pi / 2) / nphi
    R, P, T = np.meshgrid(r, ph, th, indexing="ij")
    f = np.sin(9 * R + 0.3) * np.sin(3 * T) * np.cos(4 * P)
    fr = 9 * np.cos(9 * R + 0.3) * np.sin(3 * T) * np.cos(4 * P)
    ft_ = 3 * np.sin(9 * R + 0.3) * np.cos(3 * T) * np.cos(4 * P)
    fp = -4 * np.sin(9 * R + 0.3) * np.sin(3 * T) * np.sin(4 * P)
    dri = (nr - 1) / (r[-1] - r[0])
    dti = (nth - 1) / (th[-1] - th[0])
    dpi = nphi / (np.pi / 2)
    ops = O.Operators(nr, nth, nphi, dri, dti, dpi)
    x = torch.as_tensor(f, device="cuda")
    gr = ops.dbydr(x, 1).cpu().numpy()
    gt = ops.dbydt(x, 1).cpu().numpy()
    gp = ops.dbydp(x).cpu().numpy()
    assert np.abs(gr - fr).max() < 2e-6 * np.abs(fr).max()
    assert np.abs(gt - ft_).max() < 2e-6 * np.abs(ft_).max()
    assert np.abs(gp - fp).max() < 1e-8 * np.abs(fp).max()
    ops.close()


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
