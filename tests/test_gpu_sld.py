"""GPU parity of the SLD diffusion (Compute_Diffusion, diffusion.f90:70-1495) and of the
SLDiffusion branch of Step_RK4 / Step_RK3_TVD / Step_RK2_TVD (step_rk4.f90) against the CPU
oracle, through the C ABI (cssx_diffusion, cssx_rk4, cssx_rk3_tvd, cssx_rk2_tvd).

Tolerances, relative to each field's max norm: FP64 1e-11 (the same operations in the same
order; the face tables are host-precomputed products, so a few ulps differ); FP32 states in
fluctuation form 1e-4 (limited differences of neighbouring FP32 values, eps32 ~ 6e-8 times
the ratio of a field to its neighbour differences).
"""
import numpy as np
import pytest
import torch

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


def _state(m, vel=1e-2, seed=0):
    rng = np.random.default_rng(seed)
    st = m.initial_state()
    shp = st.shape[1:]
    for iv in (1, 2, 3):
        st[iv] = vel * rng.standard_normal(shp)
    st[0] *= 1 + 1e-3 * rng.standard_normal(shp)
    if st.shape[0] == 8:
        st[5:8] *= 1e3
    return st


VARIANTS = {
    "default": dict(sld_coef=0.3),
    "flags": dict(sld_coef=0.2, Mass_Diffusion=False, dstype=0, sld_pr=0.7, sld_prm=1.3,
                  Top_Perfect_Cond=True, Bottom_Perfect_Cond=True, Theta_Perf_Conductor=True,
                  Theta_Normal_Field=False, Do_Viscous_Heating=False),
    "hydro": dict(sld_coef=0.4, magnetic=False, Top_Stress_Free=False, Bottom_Stress_Free=False),
    # SLD beside the explicit diffusion: constant-flux entropy ghosts from the wall luminosity
    # (diffusion.f90:614-618, :688-692)
    "explicit": dict(sld_coef=0.3, Laplacian=True, kappa_value=1e-3, lumr1=0.3, lumr2=0.2),
    "explicit_top": dict(sld_coef=0.3, SLaplacian=True, kappa_value=1e-3, lumr1=0.3, lumr2=0.2,
                         Top_Constant_Temp=False, Top_Constant_Flux=True),
}
SHAPES = [(24, 40, 32), (40, 36, 72)]      # (nr, nth, nphi)


@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("shape", SHAPES)
def test_diffusion_matches_oracle(mhd, shape, variant):
    nr, nth, nphi = shape
    print(f"diffusion parity {variant} {shape}", flush=True)
    m = Model(Config(nr=nr, nth=nth, nphi=nphi, **VARIANTS[variant]))
    st = _state(m)
    ref = oracle.diffusion(m, st)
    sol = mhd(m)
    x = torch.as_tensor(st, device="cuda")
    got = sol.diffusion(x).cpu().numpy()
    e = _rel_fields(got, ref)
    for iv in range(st.shape[0]):
        if np.abs(ref[iv]).max() == 0:
            assert np.all(got[iv] == 0), iv
    assert e.max() < 1e-11, e
    got32 = sol.diffusion(sol.to_f32(x)).double().cpu().numpy()
    e32 = _rel_fields(got32, ref)
    print("f32 rel err per field", e32)
    assert e32.max() < 1e-4, e32
    sol.close()


# theta lengths over several 64-column wavefronts: 65 and 129 put column nth-2 (whose stencil
# reads the upper ghost) on the last lane of the tile before the last, 130 spans three tiles
@pytest.mark.parametrize("variant", ["default", "flags"])
@pytest.mark.parametrize("shape", [(10, 65, 12), (8, 129, 10), (9, 130, 8)])
def test_diffusion_theta_tiles_match_oracle(mhd, shape, variant):
    nr, nth, nphi = shape
    print(f"diffusion theta tiles {variant} {shape}", flush=True)
    m = Model(Config(nr=nr, nth=nth, nphi=nphi, **VARIANTS[variant]))
    st = _state(m, seed=7)
    ref = oracle.diffusion(m, st)
    sol = mhd(m)
    x = torch.as_tensor(st, device="cuda")
    got = sol.diffusion(x).cpu().numpy()
    e = _rel_fields(got, ref)
    assert e.max() < 1e-11, e
    got32 = sol.diffusion(sol.to_f32(x)).double().cpu().numpy()
    assert _rel_fields(got32, ref).max() < 1e-4
    sol.close()


def test_diffusion_given_b_equals_computed(mhd):
    m = Model(Config(nr=20, nth=24, nphi=16, sld_coef=0.3))
    st = _state(m, seed=2)
    sol = mhd(m)
    x = torch.as_tensor(st, device="cuda")
    _, B = sol.rhs(x, want_b=True)
    a = sol.diffusion(x)
    b = sol.diffusion(x, b=B)
    assert torch.equal(a, b)
    sol.close()


@pytest.mark.parametrize("scheme", ["rk4", "rk3_tvd", "rk2_tvd"])
@pytest.mark.parametrize("variant", ["default", "hydro"])
def test_step_sld_matches_oracle(mhd, variant, scheme):
    print(f"{scheme} sld parity {variant}", flush=True)
    m = Model(Config(nr=24, nth=40, nphi=32, omega0=0.3, **VARIANTS[variant]))
    st = oracle.bound(m, _state(m, seed=4))
    dt, n = 2e-4, 2
    ref = oracle.rk4(m, st, dt, n, scheme)
    sol = mhd(m)
    y = torch.as_tensor(st, device="cuda").contiguous()
    sol.rk4(y, dt, n, "f64", scheme)
    got = y.cpu().numpy()
    assert _rel_fields(got, ref).max() < 1e-12
    d0 = np.abs(ref - st).max(axis=(1, 2, 3))
    chg = np.abs((got - st) - (ref - st)).max(axis=(1, 2, 3))
    assert np.all(chg <= 1e-9 * d0 + 1e-300), (chg, d0)
    y32 = torch.as_tensor(st, device="cuda").contiguous()
    sol.rk4(y32, dt, n, "f32", scheme)
    chg32 = np.abs((y32.cpu().numpy() - st) - (ref - st)).max(axis=(1, 2, 3))
    print("f32 change err / change", chg32 / np.maximum(d0, 1e-300))
    assert np.all(chg32 <= 1e-3 * d0 + 1e-300), (chg32, d0)
    sol.close()


RK4_VARIANTS = dict(VARIANTS, nomass=dict(sld_coef=0.3, Mass_Diffusion=False))


@pytest.mark.parametrize("variant", ["explicit", "flags", "nomass"])
def test_rk4_sld_variants_match_oracle(mhd, variant):
    """RK4 stage 1 applies the SLD viscous-heating and mass-diffusion corrections in its final
    kernel: beside the explicit diffusion (Laplacian, whose terms share that kernel), with both
    corrections off (the rho..s march then keeps its single pass), and with viscous heating
    alone (rho undiffused: the march stores its zero tendency)."""
    print(f"rk4 sld parity {variant}", flush=True)
    m = Model(Config(nr=24, nth=40, nphi=32, omega0=0.3, **RK4_VARIANTS[variant]))
    st = oracle.bound(m, _state(m, seed=5))
    # the explicit diffusion (kappa 1e-3) limits the stable step at this grid
    dt, n = (1e-5 if variant == "explicit" else 1e-4), 2
    ref = oracle.rk4(m, st, dt, n, "rk4")
    assert np.isfinite(ref).all()
    sol = mhd(m)
    y = torch.as_tensor(st, device="cuda").contiguous()
    sol.rk4(y, dt, n, "f64", "rk4")
    got = y.cpu().numpy()
    d0 = np.abs(ref - st).max(axis=(1, 2, 3))
    chg = np.abs((got - st) - (ref - st)).max(axis=(1, 2, 3))
    # the explicit terms' FP64 parity is 1e-9 of the change (their own RK4 test's bound)
    assert np.all(chg <= 1e-9 * d0 + 1e-300), (chg, d0)
    y32 = torch.as_tensor(st, device="cuda").contiguous()
    sol.rk4(y32, dt, n, "f32", "rk4")
    chg32 = np.abs((y32.cpu().numpy() - st) - (ref - st)).max(axis=(1, 2, 3))
    assert np.all(chg32 <= 1e-3 * d0 + 1e-300), (chg32, d0)
    sol.close()


def test_c2_shape_sld_step_fp32_close(mhd):
    """Benchmark shape 128x256x512 with SLD on: FP32 and FP64 two-step states agree."""
    c = Config(nr=128, nth=256, nphi=512, sld_coef=0.1)
    m = Model(c)
    print("c2 sld fp32-vs-fp64", flush=True)
    st = torch.as_tensor(m.initial_state(), device="cuda")
    sol = mhd(m)
    dt = sol.cfl(st)
    y64 = st.clone()
    sol.rk4(y64, dt, 2, "f64")
    y32 = st.clone()
    sol.rk4(y32, dt, 2, "f32")
    torch.cuda.synchronize()
    assert torch.isfinite(y64).all() and torch.isfinite(y32).all()
    for iv in range(8):
        ch = (y64[iv] - st[iv]).abs().max().item()
        err = (y32[iv] - y64[iv]).abs().max().item()
        sc = y64[iv].abs().max().item()
        assert err <= 1e-5 * sc + 1e-3 * ch, (iv, err, sc, ch)
    sol.close()
