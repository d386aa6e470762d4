"""BF16 MFMA network kernels against the float64 numpy restatement
(oracle/surrogate_ref.py, itself pinned by finite differences).

Inputs and weights are drawn on the bf16 grid so both sides see the same
operands; the kernels accumulate in FP32, so FP32 outputs agree to ~1e-5
relative and bf16 outputs to the bf16 rounding of the result (2^-8)."""
import numpy as np
import pytest

from muphfasa_amd import surrogate as S
from oracle import surrogate_ref as R

pytestmark = pytest.mark.gpu


def _bf16(torch, a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev).to(torch.bfloat16)


def _to_np(t):
    return t.float().cpu().numpy().astype(np.float64)


def _rand(rng, shape, scale=1.0):
    import torch
    a = (rng.standard_normal(shape) * scale).astype(np.float32)
    return torch.from_numpy(a).to(torch.bfloat16).float().numpy().astype(np.float64)


SHAPES = [(2, 9, 7, 8, 32, 1), (1, 16, 16, 32, 64, 2), (2, 8, 12, 64, 32, 1),
          (1, 10, 10, 128, 128, 2), (1, 6, 6, 32, 8, 1), (3, 5, 11, 16, 24, 1),
          # the warm-start network's deep layers on small maps
          (2, 4, 4, 96, 120, 2), (2, 2, 2, 120, 120, 2), (2, 1, 1, 120, 120, 1),
          (2, 2, 2, 120, 96, 1), (1, 3, 5, 120, 120, 2), (2, 16, 16, 16, 8, 1)]


@pytest.mark.parametrize("N,H,W,Cin,Cout,stride", SHAPES)
def test_conv_forward(dev, N, H, W, Cin, Cout, stride):
    import torch
    rng = np.random.default_rng(Cin * 7 + Cout)
    x = _rand(rng, (N, H, W, Cin))
    w = _rand(rng, (Cout, 3, 3, Cin), 0.2)
    b = rng.standard_normal(Cout).astype(np.float32)
    ref = R.conv3x3(x, w, b.astype(np.float64), stride)
    Ho, Wo = ref.shape[1], ref.shape[2]
    xt, wt = _bf16(torch, x, dev), _bf16(torch, w, dev)
    bt = torch.from_numpy(b).to(dev)
    yf = torch.empty((N, Ho, Wo, Cout), dtype=torch.float32, device=dev)
    y = torch.empty((N, Ho, Wo, Cout), dtype=torch.bfloat16, device=dev)
    S.conv(xt, wt, bt, y, yf, stride, relu=True)
    torch.cuda.synchronize()
    refr = np.maximum(ref, 0)
    scale = np.abs(ref).max()
    assert np.abs(_to_np(yf) - refr).max() <= 1e-5 * scale
    assert np.abs(_to_np(y) - refr).max() <= 2 ** -8 * scale


@pytest.mark.parametrize("N,H,W,Cin,Cout,stride", SHAPES)
def test_conv_backward(dev, N, H, W, Cin, Cout, stride):
    import torch
    rng = np.random.default_rng(Cin + 3 * Cout)
    x = np.maximum(_rand(rng, (N, H, W, Cin)), 0)          # a ReLU output (mask source)
    w = _rand(rng, (Cout, 3, 3, Cin), 0.2)
    Ho, Wo = (H + stride - 1) // stride, (W + stride - 1) // stride
    dy = _rand(rng, (N, Ho, Wo, Cout))
    dx_ref, dw_ref, db_ref = R.conv3x3_bwd(x, w, dy, stride)
    xt, wtt, dyt = _bf16(torch, x, dev), _bf16(torch, w.transpose(1, 2, 3, 0), dev), _bf16(torch, dy, dev)
    dx = torch.empty((N, H, W, Cin), dtype=torch.bfloat16, device=dev)
    S.conv_bwd_data(dyt, wtt, dx, xt, stride)
    dw = torch.zeros((Cout, 9 * Cin), dtype=torch.float32, device=dev)
    db = torch.zeros(Cout, dtype=torch.float32, device=dev)
    S.conv_bwd_weight(xt, dyt, dw, db, stride)
    torch.cuda.synchronize()
    masked = dx_ref * (x > 0)
    assert np.abs(_to_np(dx) - masked).max() <= 2 ** -7 * np.abs(dx_ref).max()
    assert np.abs(_to_np(dw) - dw_ref.reshape(Cout, -1)).max() <= 1e-5 * np.abs(dw_ref).max()
    assert np.abs(_to_np(db) - db_ref).max() <= 1e-5 * np.abs(db_ref).max()
    # accumulate mode adds onto the existing gradient
    S.conv_bwd_data(dyt, wtt, dx, xt, stride, accumulate=True)
    torch.cuda.synchronize()
    assert np.abs(_to_np(dx) - 2 * masked).max() <= 2 ** -6 * np.abs(dx_ref).max()


def _net_params_np(net):
    p = {}
    for name, l in net.layers.items():
        w = l["w"].float().cpu().numpy().astype(np.float64)           # bf16 working copy
        p[name] = (w, l["b32"].cpu().numpy().astype(np.float64))
    return p


def test_network_gradient_against_float64_reference(dev):
    """Whole-network weight gradients against float64.  The GPU rounds every
    activation and data gradient to bf16, so agreement improves with the
    number of pixels each weight gradient sums over; at 128^2 the deepest
    layers (8^2 and 4^2 maps) still sum over enough of them."""
    import torch
    H = 128
    net = S.WarmStartNet(H=H, W=H, seed=3)
    rng = np.random.default_rng(0)
    solid = np.zeros((2, H, H), np.uint8)
    solid[0, 40:64, 16:80] = 1
    re = np.array([0.3, -0.7], np.float32)
    x0 = torch.empty((2, H, H, 8), dtype=torch.bfloat16, device=dev)
    st = torch.from_numpy(solid).to(dev)
    S._ck(S.lib().mfa_nn_make_input(S._p(st), S._p(torch.from_numpy(re).to(dev)), S._p(x0), 2, H,
                                    H, S._s()))
    target = rng.standard_normal((2, H, H, 3)).astype(np.float32)
    tt = torch.from_numpy(target).to(dev)
    out = net.forward(x0)
    g = torch.empty((2, H, H, 8), dtype=torch.bfloat16, device=dev)
    part = torch.zeros((2 * H * H + 255) // 256, dtype=torch.float32, device=dev)
    count = float((solid == 0).sum() * 3)
    S._ck(S.lib().mfa_nn_mse(S._p(out), S._p(tt), S._p(st), S._p(g), S._p(part), 2 * H * H, 8, 3,
                             1.0 / count, S._s()))
    net.zero_grad()
    net.backward(g)
    torch.cuda.synchronize()
    loss_gpu = part.sum().item() / count
    params = _net_params_np(net)
    x0n = _to_np(x0)
    o, act, pre = R.forward(params, S.LAYERS, S.SKIPS, x0n)
    loss_ref, g_ref = R.mse(o, target.astype(np.float64), solid, 3)
    grads = R.backward(params, S.LAYERS, S.SKIPS, act, pre, g_ref)
    assert abs(loss_gpu - loss_ref) <= 2e-2 * loss_ref
    assert np.abs(_to_np(out) - o).max() <= 3e-2 * np.abs(o).max()
    bad = []
    for name, l in net.layers.items():
        gw = l["gw"].cpu().numpy().reshape(-1).astype(np.float64)
        rw = grads[name][0].reshape(-1)
        cos = gw @ rw / (np.linalg.norm(gw) * np.linalg.norm(rw) + 1e-300)
        rel = np.linalg.norm(gw - rw) / (np.linalg.norm(rw) + 1e-300)
        print(f"{name}: cos={cos:.5f} rel={rel:.3e}")
        if not (cos > 0.995 and rel < 0.1):
            bad.append(name)
    assert not bad, bad


def test_training_reduces_loss(dev):
    import torch
    net = S.WarmStartNet(H=32, W=32, seed=1)
    rng = np.random.default_rng(2)
    n = 8
    solid = np.zeros((n, 32, 32), np.uint8)
    re = np.linspace(-1, 1, n).astype(np.float32)
    yy, xx = np.meshgrid((np.arange(32) + 0.5) / 32, (np.arange(32) + 0.5) / 32, indexing="ij")
    target = np.stack([np.stack([0.1 * r + 0 * xx, np.sin(np.pi * yy) * (1 + r), xx * yy * r], -1)
                       for r in re]).astype(np.float32)
    st = torch.from_numpy(solid).to(dev)
    x0 = torch.empty((n, 32, 32, 8), dtype=torch.bfloat16, device=dev)
    S._ck(S.lib().mfa_nn_make_input(S._p(st), S._p(torch.from_numpy(re).to(dev)), S._p(x0), n, 32,
                                    32, S._s()))
    tt = torch.from_numpy(target).to(dev)
    g = torch.empty((n, 32, 32, 8), dtype=torch.bfloat16, device=dev)
    part = torch.zeros(4096, dtype=torch.float32, device=dev)
    count = float(n * 32 * 32 * 3)
    losses = []
    for it in range(150):
        out = net.forward(x0)
        S._ck(S.lib().mfa_nn_mse(S._p(out), S._p(tt), S._p(st), S._p(g), S._p(part), n * 32 * 32, 8,
                                 3, 1.0 / count, S._s()))
        net.zero_grad()
        net.backward(g)
        net.adam_step(2e-3)
        losses.append(part.sum().item() / count)
    print(losses[0], losses[-1])
    assert losses[-1] < 0.1 * losses[0]
