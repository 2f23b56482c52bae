"""HIP-graph capture of the solver's own RK4 step, and RK4 with the learned closure attached.

* Solver.capture records whole Step_RK4 steps (step_rk4.f90:339-493: four RHS evaluations,
  the SLD diffusion, the fused stage updates and Bound) in one graph.  Each replay must equal
  eager stepping of a copy bit for bit, replay after replay, with eager work and
  synchronizations in between.  The graph runs stage 1's SLD on the captured stream and the
  eager steps on the solver's SLD stream, so this also holds the forked SLD to the same bits.
* With a closure attached, RK4 must equal the stages composed by hand from ``rhs`` (which
  then includes the closure tendency), ``diffusion`` and ``bound`` with the RK4 weights:
  FP64 to rounding, and the FP32-bulk path (FP32 stages and accumulator, FP64 state) against
  the same composition carried out in FP32 with the fused kernel's rounding points.
* Captured with a closure attached (the closure then runs on the captured stream), replay
  equals eager stepping (which runs it on the forked side stream) bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from cssx.model import Config, Model

pytestmark = pytest.mark.gpu


def _solver(shape=(24, 32, 48), sld=0.1, seed=(2, 3)):
    from cssx.mhd import Solver
    c = Config(nr=shape[0], nth=shape[1], nphi=shape[2], sld_coef=sld, seed_s=seed[0], seed_a=seed[1])
    m = Model(c)
    sol = Solver(m)
    y = torch.as_tensor(m.initial_state(), device="cuda").contiguous()
    return sol, y


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("scheme,nsteps", [("rk4", 1), ("rk4", 2), ("rk3_tvd", 1)])
def test_captured_steps_equal_eager(precision, scheme, nsteps):
    print(f"capture {precision} {scheme} x{nsteps}", flush=True)
    sol, y = _solver()
    dt = sol.cfl(y)
    ref = y.clone()
    cap = y.clone()
    g = sol.capture(cap, dt, nsteps, precision, scheme)
    assert torch.equal(cap, y)                        # capture leaves the state untouched
    for rep in range(3):
        g.replay(1)
        sol.rk4(ref, dt, nsteps, precision, scheme)
        torch.cuda.synchronize()
        assert torch.equal(cap, ref), (rep, (cap - ref).abs().max().item())
        # eager work on the same solver between replays (its scratch is shared with the graph)
        other = y.clone()
        sol.rk4(other, dt, 1, precision, scheme)
        torch.cuda.synchronize()
    assert torch.isfinite(cap).all()
    sol.close()


def _closure(width=128, seed=21, scale=1.0):
    from cssx import closure as cl
    rng = np.random.default_rng(seed)
    net = cl.Network(width, width, max_rows=4096, seed=seed)
    xm, xs = rng.standard_normal(216) * 1e-3, rng.uniform(1e-3, 1e-2, 216)
    ym, ys = rng.standard_normal(8) * 1e-2 * scale, rng.uniform(0.1, 1.0, 8) * scale
    net.set_norm(xm, xs, ym, ys)
    return net


def _attach(sol, net):
    from cssx import _lib
    assert _lib.lib().cssx_mhd_set_closure(sol._h, net.handle if net is not None else None) == 0


def _manual_rk4_f64(sol, y0, dt):
    """Step_RK4 with SLDiffusion composed from rhs / diffusion / bound (step_rk4.f90:366-493)."""
    k1, b = sol.rhs(y0, want_b=True)
    acc = sol.diffusion(y0, b) + k1 / 6.0
    y = sol.bound((y0 + 0.5 * dt * k1).contiguous())
    k = sol.rhs(y)
    acc = acc + k / 3.0
    y = sol.bound((y0 + 0.5 * dt * k).contiguous())
    k = sol.rhs(y)
    acc = acc + k / 3.0
    y = sol.bound((y0 + dt * k).contiguous())
    k = sol.rhs(y)
    acc = acc + k / 6.0
    return sol.bound((y0 + dt * acc).contiguous())


def _manual_rk4_f32(sol, y0, dt):
    """The FP32-bulk step composed by hand: FP32 fluctuation-form stage states and accumulator
    (acc = D + k1/6, += k2/3, += k3/3, += k4/6 in FP32), the new state y0 + dt acc in FP64."""
    bg = sol.background(y0.device)
    y0f = (y0 - bg).float().contiguous()
    k1, b = sol.rhs(y0f, want_b=True)
    acc = sol.diffusion(y0f, b) + k1 / 6.0
    h = np.float32(0.5 * dt)
    y = sol.bound((y0f + h * k1).contiguous())
    k = sol.rhs(y)
    acc = acc + k / 3.0
    y = sol.bound((y0f + h * k).contiguous())
    k = sol.rhs(y)
    acc = acc + k / 3.0
    y = sol.bound((y0f + np.float32(dt) * k).contiguous())
    k = sol.rhs(y)
    acc = acc + k / 6.0
    return sol.bound((y0 + dt * acc.double()).contiguous())


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_rk4_with_closure_matches_composed_stages(precision):
    print(f"closure rk4 {precision}", flush=True)
    sol, y0 = _solver(shape=(16, 32, 32))
    net = _closure()
    _attach(sol, net)
    dt = sol.cfl(y0)
    got = sol.rk4(y0.clone(), dt, 1, precision)
    want = _manual_rk4_f64(sol, y0, dt) if precision == "f64" else _manual_rk4_f32(sol, y0, dt)
    _attach(sol, None)
    plain = sol.rk4(y0.clone(), dt, 1, precision)
    torch.cuda.synchronize()
    for iv in range(8):
        sc = y0[iv].abs().max().item()
        ch = (want[iv] - y0[iv]).abs().max().item()
        err = (got[iv] - want[iv]).abs().max().item()
        tol = 1e-12 * sc + 1e-10 * ch if precision == "f64" else 2e-7 * sc + 1e-4 * ch
        assert err <= tol, (iv, err, sc, ch)
    # the closure moved the step (so the comparison above saw it)
    moved = max(((got[iv] - plain[iv]).abs().max() / (got[iv] - y0[iv]).abs().max().clamp_min(1e-300)).item()
                for iv in range(8))
    assert moved > 1e-3, moved
    net.close()
    sol.close()


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_captured_closure_step_equals_eager(precision):
    print(f"capture with closure {precision}", flush=True)
    sol, y = _solver(shape=(16, 32, 32))
    net = _closure()
    _attach(sol, net)
    dt = sol.cfl(y)
    cap = y.clone()
    ref = y.clone()
    g = sol.capture(cap, dt, 1, precision)
    for rep in range(3):
        g.replay(1)
        sol.rk4(ref, dt, 1, precision)                # eager: the closure on the side stream
        torch.cuda.synchronize()
        assert torch.equal(cap, ref), (rep, (cap - ref).abs().max().item())
    _attach(sol, None)
    net.close()
    sol.close()
