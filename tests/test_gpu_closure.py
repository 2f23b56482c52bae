"""GPU parity of the learned-closure kernels (csrc/cssx_closure.hip) against the CPU oracle
(oracle/closure.py), and what the trained network gains.

Tolerances: GEMM FP32 accumulation of exact BF16 products, 1e-5 of the result scale (K <= 256)
and 1e-4 (K = 8192); network outputs and gradients 5e-3 / 1e-2 of their max norm (the GPU's
FP32 pre-activations can round a hidden BF16 activation the other way than the FP64 oracle);
Adam 1e-6; box filter 1e-6; raw features exact up to the FP32 cast, standardized features
within one BF16 ulp.
"""
import numpy as np
import pytest
import torch

from oracle import closure as oc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cl():
    assert torch.cuda.is_available()
    from cssx import closure
    return closure


def _bf16(x):
    return torch.as_tensor(np.asarray(x, np.float32), device="cuda").to(torch.bfloat16).contiguous()


@pytest.mark.parametrize("M,N,K,ta,tb,ks", [(16, 16, 32, 0, 0, 1), (33, 47, 216, 0, 0, 1), (128, 8, 100, 1, 0, 1),
                                            (216, 128, 8192, 1, 0, 16), (300, 128, 8, 0, 1, 1),
                                            (64, 128, 128, 0, 1, 1)])
def test_gemm_bf16_matches_fp64(cl, M, N, K, ta, tb, ks):
    from cssx import _lib
    print(f"gemm {M}x{N}x{K} ta={ta} tb={tb} ks={ks}", flush=True)
    rng = np.random.default_rng(M * 7 + N + K)
    A = oc.bf16(rng.standard_normal((M, K)).astype(np.float32))
    B = oc.bf16(rng.standard_normal((K, N)).astype(np.float32))
    At = _bf16(A.T.copy() if ta else A)
    Bt = _bf16(B.T.copy() if tb else B)
    sam, sak = (1, M) if ta else (K, 1)
    sbk, sbn = (1, K) if tb else (N, 1)
    C = torch.empty((M, N), dtype=torch.float32, device="cuda")
    rc = _lib.lib().cssx_gemm_bf16(At.data_ptr(), sam, sak, Bt.data_ptr(), sbk, sbn, M, N, K, ks, C.data_ptr(), N,
                                   None)
    assert rc == 0
    ref = A.astype(np.float64) @ B.astype(np.float64)
    err = np.abs(C.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err < (1e-4 if K > 256 else 1e-5), err


H1, H2 = 128, 128


def _net(cl, seed=11, rows=4096):
    return cl.Network(H1, H2, max_rows=rows, seed=seed)


def test_forward_and_gradients_match_oracle(cl):
    print("mlp forward/grad parity", flush=True)
    net = _net(cl)
    p = net.params().astype(np.float64)
    rng = np.random.default_rng(5)
    rows = 300
    X = oc.bf16(rng.standard_normal((rows, 216)).astype(np.float32))
    T = rng.standard_normal((rows, 8)).astype(np.float32)
    xb = _bf16(X)
    tt = torch.as_tensor(T, device="cuda")
    Y = net.forward(xb).cpu().numpy()
    Yo, _ = oc.forward(p, X, H1, H2)
    assert np.abs(Y - Yo).max() < 5e-3 * np.abs(Yo).max()
    loss = net.grad(xb, tt).item()
    g = net.grads().astype(np.float64)
    lo, go = oc.grads(p, X, T, H1, H2)
    assert abs(loss - lo) < 1e-3 * lo
    W1, b1, W2, b2, W3, b3 = oc.unpack(g, H1, H2)
    V1, c1, V2, c2, V3, c3 = oc.unpack(go, H1, H2)
    for a, b, name in ((W1, V1, "W1"), (b1, c1, "b1"), (W2, V2, "W2"), (b2, c2, "b2"), (W3, V3, "W3"),
                       (b3, c3, "b3")):
        assert np.abs(a - b).max() < 1e-2 * np.abs(b).max(), name
    net.close()


def test_train_step_is_adam_on_the_gpu_gradient(cl):
    print("mlp adam step", flush=True)
    net = _net(cl, seed=12)
    p0 = net.params().astype(np.float64)
    rng = np.random.default_rng(6)
    xb = _bf16(rng.standard_normal((256, 216)))
    tt = torch.as_tensor(rng.standard_normal((256, 8)).astype(np.float32), device="cuda")
    net.train_step(xb, tt, 1e-3)
    g = net.grads().astype(np.float64)
    p1 = net.params().astype(np.float64)
    pe, m, v = oc.adam(p0, g, np.zeros_like(g), np.zeros_like(g), 1, 1e-3)
    assert np.abs(p1 - pe).max() < 1e-6
    net.train_step(xb, tt, 1e-3)
    g2 = net.grads().astype(np.float64)
    pe2, _, _ = oc.adam(pe, g2, m, v, 2, 1e-3)
    assert np.abs(net.params() - pe2).max() < 2e-6
    net.close()


def test_training_fits_a_teacher_network(cl):
    """A student trained by the GPU kernels fits a random teacher network's outputs."""
    print("mlp teacher fit", flush=True)
    rng = np.random.default_rng(7)
    n = 32768
    X = oc.bf16(rng.standard_normal((n, 216)).astype(np.float32))
    teacher = _net(cl, seed=99).params().astype(np.float64)
    Yt, _ = oc.forward(teacher, X, H1, H2, exact=True)
    Yt = ((Yt - Yt.mean(0)) / Yt.std(0)).astype(np.float32)
    net = _net(cl, seed=13, rows=1024)
    net.set_norm(np.zeros(216), np.ones(216), np.zeros(8), np.ones(8))
    xs = _bf16(X)
    ts = torch.as_tensor(Yt, device="cuda")
    losses = []
    for ep in range(30):
        perm = torch.as_tensor(rng.permutation(n), device="cuda")
        acc = 0.0
        for b in range(n // 1024):
            idx = perm[b * 1024:(b + 1) * 1024]
            acc += net.train_step(cl.gather_rows(xs, idx), cl.gather_rows(ts, idx), 2e-3).item()
        losses.append(acc / (n // 1024))
    assert losses[-1] < 0.25 * losses[0], losses
    net.close()


@pytest.mark.parametrize("prec", [torch.float64, torch.float32])
def test_box_filter_and_features_match_oracle(cl, prec):
    from cssx.model import Config, Model
    print(f"box filter / features {prec}", flush=True)
    rng = np.random.default_rng(8)
    st = rng.standard_normal((8, 16, 24, 32))
    x = torch.as_tensor(st, dtype=prec, device="cuda")
    got = cl.box_filter(x, 4).cpu().numpy()
    ref = oc.box_filter(x.double().cpu().numpy(), 4)
    assert np.abs(got - ref).max() < 1e-6 * np.abs(ref).max()
    m = Model(Config(nr=16, nth=32, nphi=24))
    raw = cl.features(x, m).cpu().numpy()
    fo = oc.features(x.double().cpu().numpy(), m.rho_bar, m.s_bar)
    assert np.abs(raw - fo).max() <= 2e-7 * np.abs(fo).max()
    raw_fl = cl.features(x, m, fluct=True).cpu().numpy()
    assert np.abs(raw_fl - oc.features(x.double().cpu().numpy(), fluct=True)).max() <= 2e-7 * np.abs(fo).max()
    net = _net(cl, rows=64)
    mean = rng.standard_normal(216) * 0.1
    sd = rng.uniform(0.5, 2.0, 216)
    net.set_norm(mean, sd, np.zeros(8), np.ones(8))
    xb = cl.features(x, m, net=net, row0=100, rows=700).float().cpu().numpy()
    z = ((fo[100:800].astype(np.float32) - mean.astype(np.float32)) * (1.0 / sd.astype(np.float32))).astype(np.float32)
    zb = oc.bf16(z)
    assert np.all(np.abs(xb - zb) <= 2.0 ** -7 * np.abs(zb) + 1e-30)
    assert np.mean(xb == zb) > 0.99
    net.close()


@pytest.mark.parametrize("nth", [16, 20])
@pytest.mark.parametrize("precision,width", [("f64", 128), ("f32", 128), ("f64", 64), ("f32", 64)])
def test_attached_closure_adds_the_network_tendency(cl, precision, width, nth):
    """RHS with the closure attached = RHS + y_std * net(features(state)) + y_mean.  Width 128
    runs the fused register-chained inference kernel, checked against the layer-by-layer GEMM
    path (the same BF16 rounding points; a hidden activation may round the other way, so the
    bound is on the mean and a loose max); width 64 runs the layer-by-layer path itself.  nth 16
    takes the footprint-staged kernel (tiles of 16 theta points in one row), nth 20 the
    per-K-step gather kernel (tiles that cross rows)."""
    from cssx import _lib
    from cssx.mhd import Solver
    from cssx.model import Config, Model
    print(f"attached closure {precision} width {width} nth {nth}", flush=True)
    m = Model(Config(nr=16, nth=nth, nphi=32, sld_coef=0.1))
    sol = Solver(m)
    y = torch.as_tensor(m.initial_state(), device="cuda")
    st = y if precision == "f64" else sol.to_f32(y)
    rng = np.random.default_rng(9)
    net = cl.Network(width, width, max_rows=2048, seed=21)     # several chunks over the points
    xm, xs = rng.standard_normal(216) * 1e-3, rng.uniform(1e-3, 1e-2, 216)
    ym, ys = rng.standard_normal(8) * 1e-2, rng.uniform(0.1, 1.0, 8)
    net.set_norm(xm, xs, ym, ys)
    base = sol.rhs(st)
    assert _lib.lib().cssx_mhd_set_closure(sol._h, net.handle) == 0
    withc = sol.rhs(st)
    assert _lib.lib().cssx_mhd_set_closure(sol._h, None) == 0
    again = sol.rhs(st)
    assert torch.equal(again, base)
    xb = cl.features(st, m, fluct=(precision == "f32"), net=net)
    pred = torch.cat([net.forward(xb[r:r + 2048]) for r in range(0, xb.shape[0], 2048)])
    ext = (pred * torch.as_tensor(ys, dtype=torch.float32, device="cuda")
           + torch.as_tensor(ym, dtype=torch.float32, device="cuda")).t().reshape(base.shape)
    d = ((withc - base).double() - ext.double()).abs()
    scale = ext.abs().max().item()
    # the closure tendency is FP32 (its scale-and-shift may round differently from torch's)
    tol = 1e-6 * (base.abs().max().item() if precision == "f32" else 0.0) + 2e-7 * scale
    if width == 128:
        assert d.mean().item() < 1e-4 * scale and d.max().item() < 1e-2 * scale + tol
    else:
        assert d.max().item() < tol
    net.close()
    sol.close()


def test_closure_gains_over_the_plain_coarse_rhs(cl):
    """Trained on filtered fine RHS residuals of a 32x64x128 run (coarse 8x16x32, dt 1e-4,
    snapshots every 5 steps), the network, gated per field on one held-out snapshot, removes
    most of another held-out snapshot's residual and makes no field worse (a-priori gain).
    (In this early-time regime the residual is the commutation error of filtering and the
    compact derivatives on the seeded perturbations, a function of the resolved stencil.)"""
    from cssx.mhd import Solver
    from cssx.model import Config, Model
    print("closure a-priori gain", flush=True)
    fc = Config(nr=32, nth=64, nphi=128, sld_coef=0.1)
    fine = Solver(Model(fc))
    coarse = Solver(Model(cl.coarse_config(fc, 4)))
    y = torch.as_tensor(fine.model.initial_state(), device="cuda").contiguous()
    dt = 1e-4
    snaps = []
    for k in range(8):
        if k:
            fine.rk4(y, dt, 5, "f64")
        snaps.append(cl.snapshot(fine, coarse, y, 4, k * 5 * dt))
    train_ids, val_id, test_id = [0, 1, 2, 4, 5, 7], 3, 6
    ts = cl.build_training_set([snaps[i] for i in train_ids], len(train_ids) * 4096)
    net = cl.Network(128, 128, max_rows=4096, seed=7)
    hist = cl.train(net, ts, epochs=40, batch=1024, lr=1e-3)
    assert hist[-1] < 0.5 * hist[0], hist
    mask = cl.field_gate(net, snaps[val_id], ts)
    before, after = cl.apriori_gain(net, snaps[test_id], ts, mask)
    print("gate", mask, "residual rms before", before, "after", after, flush=True)
    ratio = after / before
    assert np.all(np.isfinite(ratio))
    assert np.all(ratio < 1.02), ratio
    ys = ts.y_std
    total = np.sum((after / ys) ** 2) / np.sum((before / ys) ** 2)
    assert total < 0.5, (total, ratio)
    net.close()
    fine.close()
    coarse.close()
