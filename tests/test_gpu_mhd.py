"""GPU parity of the MHD right-hand side, Bound, RK4 and CFL against the CPU oracle.

Small shapes (the oracle finishes in seconds) with deliberately uneven sizes; the full
benchmark shape is covered by size-independent properties (finite state, wall invariants,
FP32-vs-FP64 agreement).  Tolerances: FP64 1e-10 of each field's max norm (1e-12 where no cancellation); FP32 2e-5 (FP32 states in fluctuation form).
"""
import numpy as np
import pytest
import torch

import mms
import oracle
from cssx.model import Config, Model

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mhd():
    from cssx.mhd import Solver
    return Solver


def _rel_fields(a, b):
    out = []
    for iv in range(a.shape[0]):
        sc = max(np.abs(b[iv]).max(), 1e-300)
        out.append(np.abs(a[iv] - b[iv]).max() / sc)
    return np.array(out)


SHAPES = [(24, 40, 32), (40, 36, 72)]      # (nr, nth, nphi)
# theta pencils over many lanes of the contiguous sweep with a ragged last lane (129 = 16*8 + 1
# and 300 = 37*8 + 4 rows): the density closure's theta-wall values are copied out of the
# lanes that hold a pencil's first and last rows
THETA_SHAPES = [(10, 129, 12), (8, 300, 10)]


@pytest.mark.parametrize("shape", SHAPES + THETA_SHAPES)
def test_rhs_matches_oracle(mhd, shape):
    nr, nth, nphi = shape
    c = Config(nr=nr, nth=nth, nphi=nphi, omega0=0.3, damping_time=50.0)
    m = Model(c)
    st, ex, terms, Bex = mms.evaluate(m)
    ref, Bref = oracle.rhs(m, st, want_b=True)
    sol = mhd(m)
    x = torch.as_tensor(st, device="cuda")
    got, B = sol.rhs(x, want_b=True)
    e = _rel_fields(got.cpu().numpy(), ref)
    # entropy's tendency is ~1e-2 of s itself: its FP64 rounding relative to max|ds| is ~1e-11
    assert e.max() < 1e-10, e
    eb = _rel_fields(B.cpu().numpy(), Bref)
    assert eb.max() < 1e-11, eb
    # The manufactured fields are O(1) departures from the background (rho in [0.1, 1.9]), so
    # the FP32 fluctuation carries no headroom: derivative rounding eps32*|f|*dri*max|k'h|
    # (~6e-8 * 1 * 70 * 2) against tendencies as small as 1e-2 (entropy) gives up to ~1e-3.
    got32 = sol.rhs(sol.to_f32(x)).double().cpu().numpy()
    e32 = _rel_fields(got32, ref)
    assert e32.max() < 5e-3, e32
    sol.close()


@pytest.mark.parametrize("shape", SHAPES)
def test_rhs_f32_near_background(mhd, shape):
    """FP32 fluctuation form in the regime it is built for: the C2 initial state (background
    polytrope + 1e-3 entropy noise + small A) plus a 1e-2 velocity field."""
    nr, nth, nphi = shape
    c = Config(nr=nr, nth=nth, nphi=nphi, omega0=0.3)
    m = Model(c)
    st = m.initial_state()
    rng = np.random.default_rng(11)
    R, P, T = np.meshgrid(m.r, m.phi, m.theta, indexing="ij")
    st[1] += 1e-2 * np.sin(3 * R) * np.sin(2 * T) * np.cos(4 * P)
    st[2] += 1e-2 * np.cos(2 * R) * np.sin(T) * np.sin(4 * P + 0.3)
    st[3] += 1e-2 * np.sin(5 * R) * np.sin(3 * T) * np.cos(4 * P + 1.0)
    st[4] *= 1 + 1e-4 * rng.standard_normal(st[4].shape)
    ref = oracle.rhs(m, st)
    sol = mhd(m)
    x = torch.as_tensor(st, device="cuda")
    got64 = sol.rhs(x).cpu().numpy()
    assert _rel_fields(got64, ref).max() < 1e-10
    got32 = sol.rhs(sol.to_f32(x)).double().cpu().numpy()
    e32 = _rel_fields(got32, ref)
    # what rounding the input alone to the FP32 fluctuation form does to the exact RHS: the
    # density tendency near the stretched bottom wall carries rho_bar (20x the top value) times
    # the one-sided wall derivative of u_r, so it is ~1e-4 sensitive to that rounding
    bg = sol.background("cpu").numpy()
    xr = (st - bg).astype(np.float32).astype(np.float64) + bg
    cond = _rel_fields(oracle.rhs(m, xr), ref)
    print("f32 near-background rel err per field", e32, "input-rounding sensitivity", cond)
    assert e32[0] < 20 * cond[0] + 1e-6, (e32, cond)
    # u_r: gravity and the radial pressure gradient are each ~g = GM/r^2 and cancel to the
    # tendency; FP32 arithmetic leaves eps32 * g per term
    eps32 = float(np.finfo(np.float32).eps)
    assert e32[3] < 64 * eps32 * m.gravity.max() / np.abs(ref[3]).max(), e32
    others = [1, 2, 4, 5, 6, 7]
    assert e32[others].max() < 2e-6, e32
    sol.close()


def test_rhs_hydro_and_rest_state(mhd):
    c = Config(nr=32, nth=24, nphi=16, magnetic=False, damping_time=1e9)
    m = Model(c)
    st = np.zeros((5, c.nr, c.nphi, c.nth))
    st[0] = m.rho_bar[:, None, None]
    st[4] = m.s_bar[:, None, None] * (1 + 1e-3 * np.random.default_rng(3).standard_normal(st[4].shape))
    ref = oracle.rhs(m, st)
    sol = mhd(m)
    got = sol.rhs(torch.as_tensor(st, device="cuda")).cpu().numpy()
    assert _rel_fields(got, ref).max() < 1e-11
    sol.close()


def test_bound_matches_oracle(mhd):
    c = Config(nr=16, nth=12, nphi=8)
    m = Model(c)
    st = np.random.default_rng(5).standard_normal((8, 16, 8, 12))
    ref = oracle.bound(m, st)
    sol = mhd(m)
    x = torch.as_tensor(st, device="cuda")
    sol.bound(x)
    assert np.array_equal(x.cpu().numpy(), ref)
    x32 = sol.to_f32(torch.as_tensor(st, device="cuda"))
    sol.bound(x32)
    got = sol.from_f32(x32).cpu().numpy()
    walls = (ref != st)
    assert np.all(got[walls] == ref[walls])
    sol.close()


@pytest.mark.parametrize("shape", SHAPES[:1])
def test_rk4_matches_oracle(mhd, shape):
    nr, nth, nphi = shape
    c = Config(nr=nr, nth=nth, nphi=nphi, omega0=0.3, damping_time=50.0)
    m = Model(c)
    st, ex, terms, Bex = mms.evaluate(m, amp=0.2)
    st = oracle.bound(m, st)
    dt, n = 2e-4, 3
    ref = oracle.rk4(m, st, dt, n)
    sol = mhd(m)
    y = torch.as_tensor(st, device="cuda").contiguous()
    sol.rk4(y, dt, n, "f64")
    d0 = np.abs(ref - st).max()
    e = _rel_fields(y.cpu().numpy(), ref)
    assert e.max() < 1e-12, e
    # the step really moved the state: compare the change, not only the state
    chg_gpu = y.cpu().numpy() - st
    assert np.abs(chg_gpu - (ref - st)).max() < 1e-9 * d0
    y32 = torch.as_tensor(st, device="cuda").contiguous()
    sol.rk4(y32, dt, n, "f32")
    chg32 = y32.cpu().numpy() - st
    assert np.abs(chg32 - (ref - st)).max() < 1e-3 * d0
    sol.close()


@pytest.mark.parametrize("case", ["plain", "flow_rotation", "laplacian"])
def test_cfl_step(mhd, case):
    """cssx_cfl against the oracle's Compute_Step_Size (rhs_mpi.f90:92-275), which the CPU tests
    pin by closed forms; stretched grid, perturbed state, B from the seeded vector potential."""
    kw = {"plain": {}, "flow_rotation": {"omega0": 0.3, "xi_cs": 0.6},
          "laplacian": {"Laplacian": True, "SLDiffusion": False, "kappa_value": 3e-3}}[case]
    c = Config(nr=24, nth=20, nphi=16, bamp=1.0, **kw)
    m = Model(c)
    st = m.initial_state()
    if case == "flow_rotation":
        rng = np.random.default_rng(9)
        st[1:4] = 0.2 * rng.standard_normal(st[1:4].shape)
    sol = mhd(m)
    dt = sol.cfl(torch.as_tensor(st, device="cuda"))
    ref = oracle.step_size(m, st)
    assert abs(dt - ref["dt"]) < 1e-12 * ref["dt"], (dt, ref)
    if case == "plain":
        assert ref["dtmaxb"] < 0.999 * ref["dtmaxc"]      # the Alfven term decides this case
    if case == "laplacian":
        assert ref["dtmaxd"] < min(ref["dtmaxb"], ref["dtmaxc"])
    sol.close()


def test_c2_shape_step_finite_and_fp32_close(mhd):
    """Benchmark shape 128x256x512: two RK4 steps in FP64 and FP32 from the same state."""
    c = Config(**{"nr": 128, "nth": 256, "nphi": 512})
    m = Model(c)
    st = torch.as_tensor(m.initial_state(), device="cuda")
    sol = mhd(m)
    dt = sol.cfl(st)
    y64 = st.clone()
    sol.rk4(y64, dt, 2, "f64")
    y32 = st.clone()
    sol.rk4(y32, dt, 2, "f32")
    torch.cuda.synchronize()
    assert torch.isfinite(y64).all()
    for iv in range(8):
        ch = (y64[iv] - st[iv]).abs().max().item()
        err = (y32[iv] - y64[iv]).abs().max().item()
        sc = y64[iv].abs().max().item()
        assert err <= 1e-5 * sc + 1e-3 * ch, (iv, err, sc, ch)
    sol.close()


def test_device_initial_state_matches_host():
    """The GPU-built initial state (cssx_normal_fill + the same recipe) against numpy's, on an
    uneven shape: equal to the rounding of the device sin/cos/log (1e-14 of each field)."""
    m = Model(Config(nr=24, nth=40, nphi=33, seed_s=9, seed_a=10))
    host = m.initial_state()
    dev = m.initial_state(device="cuda").cpu().numpy()
    for iv in range(host.shape[0]):
        sc = max(np.abs(host[iv]).max(), 1e-300)
        assert np.abs(dev[iv] - host[iv]).max() <= 1e-14 * sc, iv
    # the perturbations themselves (not just the background) agree
    pert = host[4] - m.s_bar[:, None, None]
    assert np.abs((dev[4] - m.s_bar[:, None, None]) - pert).max() <= 1e-9 * np.abs(pert).max()


def test_c4_shape_two_steps_fp32_within_1e5(mhd):
    """The 256 x 512 x 1024 workload at full size: two RK4 steps (SLD on) in FP64 and in the
    FP32-bulk path from the same state; finite, and every field of the FP32 state within 1e-5
    of the FP64 state's max norm."""
    print("c4 two steps", flush=True)
    c = Config(nr=256, nth=512, nphi=1024, sld_coef=0.1, seed_s=5, seed_a=6)
    m = Model(c)
    st = m.initial_state(device="cuda")
    sol = mhd(m)
    dt = sol.cfl(st)
    y64 = st.clone()
    sol.rk4(y64, dt, 2, "f64")
    y32 = st.clone()
    sol.rk4(y32, dt, 2, "f32")
    torch.cuda.synchronize()
    assert torch.isfinite(y64).all() and torch.isfinite(y32).all()
    for iv in range(8):
        err = (y32[iv] - y64[iv]).abs().max().item()
        sc = y64[iv].abs().max().item()
        print(f"c4 field {iv}: max|y32-y64|/max|y64| = {err / sc:.2e}", flush=True)
        assert err <= 1e-5 * sc, (iv, err, sc)
    sol.close()
    del st, y64, y32
    torch.cuda.empty_cache()


def test_conservation_totals():
    """E, L, m, mrsint of Compute_Step_Size (rhs_mpi.f90:156-183) against the oracle, which the
    CPU tests pin by closed forms, and m of a uniform density against the exact shell-sector
    volume."""
    from cssx.mhd import Solver
    from cssx.model import Config, Model
    print("conservation totals", flush=True)
    c = Config(nr=24, nth=32, nphi=16, omega0=0.3)
    m = Model(c)
    sol = Solver(m)
    st = m.initial_state()
    rng = np.random.default_rng(4)
    st[1:4] = 1e-2 * rng.standard_normal(st[1:4].shape)
    y = torch.as_tensor(st, device="cuda")
    got = sol.conservation(y)
    ref = oracle.step_size(m, st)
    for k in ("E", "L", "m", "mrsint"):
        assert abs(got[k] - ref[k]) < 1e-12 * abs(ref[k]), (k, got[k], ref[k])
    y1 = torch.ones_like(y)
    y1[1:4] = 0.0
    vol = (c.r2 ** 3 - c.r1 ** 3) / 3 * (np.cos(c.th1) - np.cos(c.th2)) * (c.phi2 - c.phi1)
    assert abs(sol.conservation(y1)["m"] - vol) < 0.06 * vol   # end-point quadrature: ~3.5%
    sol.close()


def test_scalar_quantities():
    """Scalar_Output's sums (report_mpi.f90:195-244) against numpy on the same state and B."""
    from cssx import diagnostics as dg
    from cssx.mhd import Solver
    from cssx.model import Config, Model
    print("scalar quantities", flush=True)
    m = Model(Config(nr=16, nth=24, nphi=16))
    sol = Solver(m)
    st = m.initial_state()
    rng = np.random.default_rng(5)
    st[1:4] = 1e-2 * rng.standard_normal(st[1:4].shape)
    y = torch.as_tensor(st, device="cuda")
    q = dg.scalar_quantities(sol, y)
    _, B = sol.rhs(y, want_b=True)
    B = B.cpu().numpy()
    rho, w, v, u, s = st[:5]
    N = rho.size
    usq = w * w + v * v + u * u
    bsq = (B ** 2).sum(0)
    tsq = rho ** m.gamm1 * np.exp(s / m.Cv) * m.invtconst * m.gamm1 * m.c.Cp
    ref = {"kinener": (rho * usq).sum() / (2 * N), "cke": (rho * u * u).sum() / (2 * N),
           "mcke": ((rho * u * u).sum() + (rho * w * w).sum()) / (2 * N), "drke": (rho * v * v).sum() / (2 * N),
           "magener": bsq.sum() / (8 * np.pi * N), "bmax": np.sqrt(bsq.max()) * 2 * np.sqrt(2 * np.pi),
           "umach": np.sqrt((usq / (2 * bsq / rho + usq + tsq)).max())}
    for k, r in ref.items():
        assert abs(q[k] - r) <= 1e-12 * abs(r), (k, q[k], r)
    sol.close()
