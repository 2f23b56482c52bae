"""Batched uneven patches (one persistent launch for all three derivatives of every patch) against
the oracle, the seeded patch generator against numpy, and the HIP-graph ensemble step against
eager stepping.  Tolerances as the derivative tests: FP64 1e-12, FP32 3e-5 of the max norm."""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu


def _smooth(a, axis):
    lo = np.concatenate([a.take([0], axis), a.take(range(a.shape[axis] - 1), axis)], axis)
    hi = np.concatenate([a.take(range(1, a.shape[axis]), axis), a.take([a.shape[axis] - 1], axis)], axis)
    return 0.25 * (lo + hi) + 0.5 * a


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_patch_fill_matches_numpy(prec):
    from cssx.model import normal
    from cssx.patches import PatchChunk
    print(f"patch fill {prec}", flush=True)
    shapes = [(9, 10, 11), (12, 8, 16)]
    ch = PatchChunk(shapes, prec=prec)
    ch.fill(17)
    for i, (nr, nth, nphi) in enumerate(shapes):
        ref = normal(17 * 1000003 + i, nr * nth * nphi).reshape(nr, nphi, nth)
        ref = _smooth(_smooth(_smooth(ref, 2), 1), 0)
        got = ch.field(i).double().cpu().numpy()
        assert np.abs(got - ref).max() < (1e-12 if prec == "f64" else 2e-6) * np.abs(ref).max()
    ch.close()


def _oracle_derivs(f, bdry):
    nr, nphi, nth = f.shape
    dr = oracle.deriv_field(0, 1, f, True, True, bdry, scale=nr - 1, hneu=nr - 1)
    dt = oracle.deriv_field(1, 1, f, True, True, bdry, scale=nth - 1, hneu=nth - 1)
    # phi with walls: the r operator on the array with phi moved to the slow axis
    fp = np.ascontiguousarray(f.transpose(1, 0, 2))
    dp = oracle.deriv_field(0, 1, fp, True, True, bdry, scale=nphi - 1, hneu=nphi - 1).transpose(1, 0, 2)
    return dr, dt, dp


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 3e-5)])
@pytest.mark.parametrize("bdry", [1, 0, 3, 2, 4])
def test_patch_derivatives_match_oracle(prec, tol, bdry):
    from cssx.patches import PatchChunk
    print(f"patches {prec} bdry={bdry}", flush=True)
    shapes = [(8, 8, 8), (33, 17, 40), (65, 24, 9), (20, 129, 31), (17, 12, 257), (40, 48, 64)]
    ch = PatchChunk(shapes, prec=prec, bdry=bdry)
    rng = np.random.default_rng(bdry)
    fs = []
    for i, (nr, nth, nphi) in enumerate(shapes):
        f = rng.standard_normal((nr, nphi, nth))
        ch.field(i).copy_(torch.as_tensor(f))
        fs.append(ch.field(i).double().cpu().numpy())
    refs = [_oracle_derivs(f, bdry) for f in fs]
    for variant in range(4):
        for t in ch.out:
            t.fill_(float("nan"))
        ch.run(variant)
        torch.cuda.synchronize()
        for i in range(len(shapes)):
            for ax in range(3):
                got = ch.field(i, ax).double().cpu().numpy()
                ref = refs[i][ax]
                err = np.abs(got - ref).max() / np.abs(ref).max()
                assert err < tol, (variant, shapes[i], ax, err)
    best, ms = ch.autotune(1)
    assert 0 <= best < 4 and all(m > 0 for m in ms)
    ch.close()


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 3e-5)])
@pytest.mark.parametrize("bdry", [1, 4])
def test_patch_long_pencils_match_oracle(prec, tol, bdry):
    """Axis lengths up to the workload's maximum of 512 on every axis: theta items of 34-64
    lanes (MR = 8), and strided r / phi items of 16-32 segments."""
    from cssx.patches import PatchChunk
    print(f"long patches {prec} bdry={bdry}", flush=True)
    shapes = [(512, 12, 10), (10, 512, 12), (12, 10, 512), (37, 300, 33), (450, 8, 17), (9, 449, 270)]
    ch = PatchChunk(shapes, prec=prec, bdry=bdry)
    rng = np.random.default_rng(100 + bdry)
    fs = []
    for i, (nr, nth, nphi) in enumerate(shapes):
        f = rng.standard_normal((nr, nphi, nth))
        ch.field(i).copy_(torch.as_tensor(f))
        fs.append(ch.field(i).double().cpu().numpy())
    refs = [_oracle_derivs(f, bdry) for f in fs]
    for variant in range(4):
        for t in ch.out:
            t.fill_(float("nan"))
        ch.run(variant)
        torch.cuda.synchronize()
        for i in range(len(shapes)):
            for ax in range(3):
                got = ch.field(i, ax).double().cpu().numpy()
                ref = refs[i][ax]
                err = np.abs(got - ref).max() / np.abs(ref).max()
                assert err < tol, (variant, shapes[i], ax, err)
    ch.close()


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_ensemble_member_in_graph_matches_member_alone(precision):
    """Members of a 48 x 96 x 96 ensemble stepped by the captured graph (four members on
    concurrent side streams) against the same members built as stand-alone solvers and stepped
    alone on the default stream (bit for bit), and, in FP64, against the oracle's Step_RK4
    (1e-11 of each field's max norm)."""
    from cssx.mhd import Solver
    from cssx.model import Model
    from cssx.patches import Ensemble
    print(f"ensemble member vs alone {precision}", flush=True)
    ens = Ensemble(members=4, shape=(48, 96, 96), seed=4, precision=precision)
    y0 = [y.clone() for y in ens.states]
    ens.capture()
    # capture steps once eagerly (warm-up) and leaves the states one step on; restart them
    for y, y_init in zip(ens.states, y0):
        y.copy_(y_init)
    ens.replay(1)
    torch.cuda.synchronize()
    for k in (0, 3):
        cfg = ens.solvers[k].model.c
        alone = Solver(Model(cfg))
        y = alone.model.initial_state(device="cuda")
        assert torch.equal(y, y0[k])
        dt = alone.cfl(y)
        assert dt == ens.dts[k]
        alone.rk4(y, dt, 1, precision)
        torch.cuda.synchronize()
        assert torch.equal(y, ens.states[k]), (k, (y - ens.states[k]).abs().max().item())
        if precision == "f64":
            ref = oracle.rk4(alone.model, y0[k].cpu().numpy(), dt, 1)
            got = ens.states[k].cpu().numpy()
            for iv in range(ref.shape[0]):
                err = np.abs(got[iv] - ref[iv]).max() / np.abs(ref[iv]).max()
                assert err < 1e-11, (k, iv, err)
        alone.close()
    ens.close()


def test_ensemble_graph_replays_match_eager():
    """Graph replays of one ensemble step (members on concurrent side streams) and eager
    multi-stream steps, each bit-identical to stepping every member on one stream."""
    from cssx.patches import Ensemble
    print("ensemble graph", flush=True)
    ens = Ensemble(members=3, shape=(16, 16, 32), seed=4)
    assert np.all((ens.ra >= 1e4) & (ens.ra <= 1e6))
    ens.capture()
    ref = [y.clone() for y in ens.states]
    for k in range(3):
        ens.replay(1)
        for sol, y, dt in zip(ens.solvers, ref, ens.dts):
            sol.rk4(y, dt, 1, ens.precision)
        torch.cuda.synchronize()
        for a, b in zip(ens.states, ref):
            assert torch.equal(a, b), k
    # eager stepping with the members spread over the side streams, against one stream
    ens.step(2)
    for sol, y, dt in zip(ens.solvers, ref, ens.dts):
        sol.rk4(y, dt, 2, ens.precision)
    torch.cuda.synchronize()
    for a, b in zip(ens.states, ref):
        assert torch.equal(a, b)
    ens.close()
