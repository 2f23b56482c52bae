"""Pins the CPU oracle's derivative operators against what the mathematics fixes.

The reference ships no golden vectors for this path; its own checks are analytic known
answers (validation.f90:17-28, CSS-Python/Unit-Tests/test_compact_diff.py:18-92).  Here:
  * factor-table fixed points (derivatives.f90:156-161: pivot -> (3+sqrt5)/6, alpha=1/3;
    (1+sqrt(1-16/121))/2, alpha=2/11),
  * polynomial exactness of every closure to its documented order, and failure one degree
    above (so a dropped or altered coefficient shows),
  * B = A.D is banded with the published interior stencil (A built independently from the
    documented tridiagonal coefficients),
  * sixth/fifth-order convergence on analytic fields, the periodic phi wrap,
  * the Python compact_fd6 restatement against the Fortran restatement where they must
    agree: dtype 0 for every ibc, and dtype 1/2/3 at n=61 (dli=60) where the Python halo
    rows' extra dli/60 factor is exactly 1.
"""
import numpy as np
import pytest

import oracle


def _poly_case(order, n, lo_wall, hi_wall, bdry, p):
    x = np.linspace(0.0, 1.0, n)
    h = x[1] - x[0]
    f = x ** p
    fp = p * x ** (p - 1) if p > 0 else 0 * x
    fpp = p * (p - 1) * x ** (p - 2) if p > 1 else 0 * x
    xl = x[0] - h * np.arange(4, 0, -1)
    xu = x[-1] + h * np.arange(1, 5)
    b1, b2 = xl ** p, xu ** p
    if lo_wall:
        b1 = np.array([fp[0], 0, 0, 0])
    if hi_wall:
        b2 = np.array([fp[-1], 0, 0, 0])
    sc = 1 / h if order == 1 else 1 / h ** 2
    d = oracle.deriv_pencil(order, f, lo_wall, hi_wall, bdry, b1, b2, scale=sc, hneu=1 / h)
    ref = fp if order == 1 else fpp
    return np.abs(d - ref).max() / max(1.0, np.abs(ref).max())


# (order, lo_wall, hi_wall, bdry) -> highest degree reproduced exactly
EXACT = {}
for _b in range(1, 5):
    EXACT[(1, 0, 0, _b)] = 6               # Lele alpha=1/3 interior + explicit 6th-order halo rows
    EXACT[(1, 1, 1, _b)] = 6 if _b == 4 else 5   # 5th-order one-sided wall closure; Neumann rows exact to 6
    EXACT[(2, 0, 0, _b)] = 7               # alpha=2/11 interior + 8th-order explicit halo rows
    EXACT[(2, 1, 1, _b)] = 6
EXACT[(1, 0, 0, 0)] = 6
EXACT[(2, 0, 0, 0)] = 7


@pytest.mark.parametrize("key", sorted(EXACT))
def test_polynomial_exactness(key):
    order, lo, hi, bdry = key
    deg = EXACT[key]
    for p in range(deg + 1):
        assert _poly_case(order, 16, lo, hi, bdry, p) < 1e-9, (key, p)
    if order == 1:
        # one degree higher is no longer exact: the documented order is the real order
        assert _poly_case(order, 16, lo, hi, bdry, deg + 1) > 1e-8, key


@pytest.mark.parametrize("order,alpha,pivot", [
    (1, 1 / 3, (3 + np.sqrt(5)) / 6),
    (2, 2 / 11, (1 + np.sqrt(1 - 16 / 121)) / 2),
])
def test_factor_fixed_point(order, alpha, pivot):
    for lo, hi, iv in [(0, 0, 0), (1, 1, 1), (1, 1, 4)]:
        g1, g2, g3 = oracle.factors(order, 64, lo, hi, iv)
        assert abs(1 / g2[32] - pivot) < 1e-15
        # gam3 = -alpha*gam2 and gam1 = -alpha*gam2(j-1) in the interior
        assert np.allclose(g3[10:50], -alpha * g2[10:50], rtol=0, atol=1e-16)
        assert np.allclose(g1[10:50], -alpha * g2[9:49], rtol=0, atol=1e-16)


def _dense_operator(order, n, lo, hi, bdry):
    """Matrix D of the pencil operator (columns = response to unit vectors), scale 1."""
    D = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        D[:, k] = oracle.deriv_pencil(order, e, lo, hi, bdry, None, None, scale=1.0, hneu=1.0)
    return D


def _lhs_matrix(order, n, lo, hi, bdry):
    """Tridiagonal A from the documented coefficients (derivatives.f90:108-125, :196-216)."""
    a = 1 / 3 if order == 1 else 2 / 11
    A = np.eye(n) + a * (np.eye(n, k=1) + np.eye(n, k=-1))
    if order == 1:
        up1 = 4.0 if bdry in (1, 3) else 0.0
        up2 = 4.0 if bdry in (1, 2) else 0.0
        r1 = (1 / 8, 3 / 4)
    else:
        dir1 = bdry in (1, 3)
        dir2 = bdry in (1, 2)
        up1 = 137 / 13 if dir1 else 707 / 193
        up2 = 137 / 13 if dir2 else 707 / 193
        r1 = (0.1, -0.35)
    if lo:
        A[0, 1] = up1
        A[1, 0], A[1, 2] = r1
    else:
        A[0, 1] = a
    if hi:
        A[-1, -2] = up2
        A[-2, -1], A[-2, -3] = r1
    else:
        A[-1, -2] = a
    return A


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("lo,hi,bdry", [(0, 0, 0), (1, 1, 1), (1, 1, 4), (1, 1, 2), (1, 1, 3)])
def test_banded_rhs(order, lo, hi, bdry):
    n = 20
    D = _dense_operator(order, n, lo, hi, bdry)
    B = _lhs_matrix(order, n, lo, hi, bdry) @ D
    # interior rows carry exactly the Lele stencil
    if order == 1:
        stencil = {-2: -1 / 36, -1: -7 / 9, 0: 0.0, 1: 7 / 9, 2: 1 / 36}
    else:
        stencil = {-2: 3 / 44, -1: 12 / 11, 0: -2 * (12 / 11 + 3 / 44), 1: 12 / 11, 2: 3 / 44}
    for j in range(2, n - 2):
        row = np.zeros(n)
        for k, c in stencil.items():
            row[j + k] = c
        assert np.abs(B[j] - row).max() < 1e-12, (j, B[j])
    if lo and order == 1:
        # wall row 2 (derivatives.f90:633-640): (-43/96, -5/6, 9/8, 1/6, -1/96)
        assert np.allclose(B[1, :5], [-43 / 96, -5 / 6, 9 / 8, 1 / 6, -1 / 96], atol=1e-12)
        assert np.abs(B[1, 5:]).max() < 1e-12


def test_convergence_sin():
    errs = []
    for n in (32, 64, 128):
        r = np.linspace(0.72, 0.97, n)
        h = r[1] - r[0]
        f = np.sin(7.0 * r)
        fp = 7.0 * np.cos(7.0 * r)
        d = oracle.deriv_pencil(1, f, 1, 1, 1, scale=1 / h)
        errs.append(np.abs(d - fp).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates.min() > 4.5, rates       # fifth-order wall closure dominates the max norm


def test_periodic_phi_wrap():
    nphi, nth, nr = 64, 3, 2
    phi = np.arange(nphi) * (np.pi / 2) / nphi
    f = np.cos(4 * phi)[None, :, None] * np.ones((nr, nphi, nth))
    d = oracle.deriv_field(2, 1, f, scale=nphi / (np.pi / 2))
    ref = (-4 * np.sin(4 * phi))[None, :, None]
    assert np.abs(d - ref).max() < 1e-8


def _py_case(n, ibc, dtype, darr_pre=True):
    rng = np.random.default_rng(100 + n + 10 * ibc + dtype)
    r1, r2 = 0.75, 1.0
    dri = 1.0 / (r2 - r1)
    h = (r2 - r1) / (n - 1)
    arr = rng.standard_normal(n)
    b1 = rng.standard_normal(4)
    b2 = rng.standard_normal(4)
    v1, v2 = rng.standard_normal(2)
    darr = np.zeros(n)
    darr[0] = v1 / dri           # Python convention: f'/dri preloaded (T4)
    darr[-1] = v2 / dri
    return arr, b1, b2, v1, v2, dri, h, darr


@pytest.mark.parametrize("ibc", [0, 1, 2, 3, 4])
def test_python_dtype0_matches_fortran_walls(ibc):
    n = 40
    arr, b1, b2, v1, v2, dri, h, darr = _py_case(n, ibc, 0)
    res, _ = oracle.compact_fd6_py(dri, arr, 0.0, 0.0, ibc, 0, darr)
    ref = oracle.deriv_pencil(1, arr, 1, 1, ibc, np.array([v1, 0, 0, 0]), np.array([v2, 0, 0, 0]),
                              scale=1 / h)
    assert np.abs(res - ref).max() < 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("dtype,lo,hi", [(3, 0, 0), (1, 0, 1), (2, 1, 0)])
@pytest.mark.parametrize("ibc", [1, 4])
def test_python_halo_rows_at_n61(dtype, lo, hi, ibc):
    n = 61        # dli = 60: the Python's dli/60 factor on the halo rows is exactly 1
    arr, b1, b2, v1, v2, dri, h, darr = _py_case(n, ibc, dtype)
    res, _ = oracle.compact_fd6_py(dri, arr, b1, b2, ibc, dtype, darr)
    fb1 = b1 if not lo else np.array([v1, 0, 0, 0])
    fb2 = b2 if not hi else np.array([v2, 0, 0, 0])
    ref = oracle.deriv_pencil(1, arr, lo, hi, ibc, fb1, fb2, scale=1 / h)
    assert np.abs(res - ref).max() < 1e-11 * np.abs(ref).max()


def test_python_halo_rows_scale_elsewhere():
    """Away from n=61 the Python's halo rows differ from the Fortran's (the dli/60 factor)."""
    n = 40
    arr, b1, b2, v1, v2, dri, h, darr = _py_case(n, 1, 3)
    res, _ = oracle.compact_fd6_py(dri, arr, b1, b2, 1, 3, darr)
    ref = oracle.deriv_pencil(1, arr, 0, 0, 1, b1, b2, scale=1 / h)
    assert np.abs(res - ref).max() > 1e-3 * np.abs(ref).max()


def test_python_ibc5_uses_n():
    n = 60        # ibc=5: dli = n = 60, halo rows exact again
    arr, b1, b2, v1, v2, dri, h, darr = _py_case(n, 5, 3)
    res, _ = oracle.compact_fd6_py(dri, arr, b1, b2, 5, 3, darr)
    ref = oracle.deriv_pencil(1, arr, 0, 0, 1, b1, b2, scale=60 * dri)
    assert np.abs(res - ref).max() < 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("ibc", [1, 4])
def test_reference_known_answer_sin(ibc):
    """CSS-Python/Unit-Tests/test_compact_diff.py:18-67 case: sin on [0.75,1], nr=128, dtype 0."""
    nr = 128
    r1, r2 = 0.75, 1.0
    dr = (r2 - r1) / (nr - 1)
    dri = 1.0 / (r2 - r1)
    r = dr * np.arange(nr) + r1
    f = np.sin(2 * np.pi * (r - r1) / (r2 - r1))
    exact = 2 * np.pi * np.cos(2 * np.pi * (r - r1) / (r2 - r1)) / (r2 - r1)
    df = None
    if ibc == 4:
        df = np.zeros(nr)
        df[0] = df[-1] = 2 * np.pi / dri / (r2 - r1)
    res, _ = oracle.compact_fd6_py(dri, f, 0.0, 0.0, ibc, 0, df)
    l2 = np.sqrt(dr * np.sum((res - exact) ** 2))
    assert l2 < 1e-6, l2


def test_c1_recipe_oracle_against_analytic_and_python():
    """The plain-derivative workload (64 x 64 x 128, tanh-stretched r, seeded 8-term field):
    the oracle's d/dr (times dx/dr), d/dtheta and d/dphi (phi halos from the field itself)
    against the analytic derivatives, with the one-sided wall closures (code 1) and with the
    exact wall derivatives as Neumann values (code 4); and the CSS-Python compact_fd6
    restatement on sampled r and theta pencils (dtype 0, where it equals the Fortran rows)."""
    import c1_field
    c = c1_field.build()
    f = c["f"]
    gr = oracle.deriv_field(0, 1, f, bdry=1, scale=c["dri"]) * c["dxdr"][:, None, None]
    gt = oracle.deriv_field(1, 1, f, bdry=1, scale=c["dti"])
    gp = oracle.deriv_field(2, 1, f, lo_wall=False, hi_wall=False, ghost_lo=c["ghost_lo"], ghost_hi=c["ghost_hi"],
                            scale=c["dpi"])

    def rel(a, b):
        return np.abs(a - b).max() / np.abs(b).max()

    # one-sided closures (3rd/4th order) dominate the wall rows; the interior is 6th order
    assert rel(gr, c["fr"]) < 5e-6 and rel(gr[4:-4], c["fr"][4:-4]) < 1e-7
    assert rel(gt, c["ft"]) < 1e-5 and rel(gt[..., 4:-4], c["ft"][..., 4:-4]) < 1e-7
    assert rel(gp, c["fp"]) < 1e-9
    # an error of one mode's sign or a dropped metric factor is O(1)
    assert rel(oracle.deriv_field(0, 1, f, bdry=1, scale=c["dri"]), c["fr"]) > 0.1
    # Neumann walls with the exact wall derivatives (slot 0 of the ghost buffers)
    nr, nphi, nth = f.shape
    gl = np.zeros((4, nphi, nth))
    gh = np.zeros((4, nphi, nth))
    gl[0] = c["fr"][0] / c["dxdr"][0]
    gh[0] = c["fr"][-1] / c["dxdr"][-1]
    g4 = oracle.deriv_field(0, 1, f, bdry=4, ghost_lo=gl, ghost_hi=gh, scale=c["dri"], hneu=c["dri"])
    assert rel(g4 * c["dxdr"][:, None, None], c["fr"]) < 2e-8
    # CSS-Python compact_fd6 (dri = 1 / domain length in index space, dli = n - 1 inside)
    rng = np.random.default_rng(11)
    for _ in range(16):
        ip, it = rng.integers(0, nphi), rng.integers(0, nth)
        pr_, _ = oracle.compact_fd6_py(1.0, f[:, ip, it], 0.0, 0.0, 1, 0)
        assert np.abs(pr_ * c["dxdr"] - gr[:, ip, it]).max() < 1e-10 * np.abs(gr).max()
        ir = rng.integers(0, nr)
        pt, _ = oracle.compact_fd6_py(1.0 / (np.pi / 2), f[ir, ip, :], 0.0, 0.0, 1, 0)
        assert np.abs(pt - gt[ir, ip, :]).max() < 1e-10 * np.abs(gt).max()
