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
          (2, 2, 2, 120, 96, 1), (1, 3, 5, 120, 120, 2), (2, 16, 16, 16, 8, 1),
          # thin layers over many pixels (the weight gradient's long reductions)
          # and sizes that leave a partial 32-pixel chunk
          (1, 64, 64, 16, 16, 1), (2, 33, 31, 8, 16, 1), (1, 37, 29, 16, 32, 2),
          # enough 64 x 64 output tiles for the forward and data gradient to
          # run without splitting k
          (8, 64, 64, 32, 64, 1)]


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


def test_fused_adam_equals_per_tensor_adam_and_weight_prep(dev):
    """mfa_nn_adam_multi over the network's flat buffers gives, bit for bit,
    what mfa_nn_adam on every weight and bias followed by
    mfa_nn_prep_weights gives: the FP32 parameters, both moments and the
    BF16 weights and their transpose, over three steps with weight decay."""
    import torch
    net = S.WarmStartNet(H=16, W=16, seed=3, dev=dev)
    rng = np.random.default_rng(5)
    ref = {k: {"w": l["w32"].clone(), "b": l["b32"].clone(),
               "m": [torch.zeros_like(l["w32"]), torch.zeros_like(l["b32"])],
               "v": [torch.zeros_like(l["w32"]), torch.zeros_like(l["b32"])],
               "wb": torch.empty_like(l["w"]), "wtb": torch.empty_like(l["wt"])}
           for k, l in net.layers.items()}
    L = S.lib()
    for t in range(1, 4):
        g = torch.from_numpy(rng.standard_normal(net.flat_g.numel()).astype(np.float32)).to(dev)
        net.flat_g.copy_(g)
        net.adam_step(1e-3, wd=1e-4)
        for k, l in net.layers.items():
            r = ref[k]
            S._ck(L.mfa_nn_adam(S._p(r["w"]), S._p(l["gw"]), S._p(r["m"][0]), S._p(r["v"][0]),
                                r["w"].numel(), 1e-3, 0.9, 0.999, 1e-8, t, 1e-4, S._s()))
            S._ck(L.mfa_nn_adam(S._p(r["b"]), S._p(l["gb"]), S._p(r["m"][1]), S._p(r["v"][1]),
                                r["b"].numel(), 1e-3, 0.9, 0.999, 1e-8, t, 0.0, S._s()))
            S._ck(L.mfa_nn_prep_weights(S._p(r["w"]), S._p(r["wb"]), S._p(r["wtb"]), l["cout"],
                                        l["cin"], S._s()))
        torch.cuda.synchronize()
        m, v = net._adam
        for k, l in net.layers.items():
            r = ref[k]
            nw = l["w32"].numel()
            assert torch.equal(l["w32"], r["w"]) and torch.equal(l["b32"], r["b"]), (t, k)
            assert torch.equal(m[l["off"]:l["off"] + nw].view_as(r["m"][0]), r["m"][0]), (t, k)
            assert torch.equal(v[l["off"] + nw:l["off"] + nw + l["cout"]], r["v"][1]), (t, k)
            assert torch.equal(l["w"], r["wb"]) and torch.equal(l["wt"], r["wtb"]), (t, k)


@pytest.mark.parametrize("C", [8, 16, 24])
def test_elementwise_kernels_vector_and_scalar_paths(dev, C):
    """upsample_add, upsample_bwd and masked_add take 8 channels per thread
    on 16-byte aligned tensors and one element per thread otherwise; both
    paths give the same bits, and match numpy (sums of bf16 values in FP32,
    rounded to bf16)."""
    import torch
    rng = np.random.default_rng(C)
    N, H, W = 2, 5, 7

    def pair(shape):
        """an aligned tensor and a misaligned view holding the same values"""
        a = _bf16(torch, rng.standard_normal(shape), dev)
        buf = torch.empty(a.numel() + 1, dtype=torch.bfloat16, device=dev)
        b = buf[1:].view(shape)
        b.copy_(a)
        return a, b

    def f(t):
        return t.float().cpu().numpy()

    x, xm = pair((N, H, W, C))
    sk, skm = pair((N, 2 * H, 2 * W, C))
    y = torch.empty((N, 2 * H, 2 * W, C), dtype=torch.bfloat16, device=dev)
    ym = torch.empty(y.numel() + 1, dtype=torch.bfloat16, device=dev)[1:].view(y.shape)
    S.upsample_add(x, sk, y)
    S.upsample_add(xm, skm, ym)
    torch.cuda.synchronize()
    up = np.repeat(np.repeat(f(x), 2, axis=1), 2, axis=2)
    ref = torch.from_numpy((up + f(sk)).astype(np.float32)).to(torch.bfloat16).float().numpy()
    assert torch.equal(y, ym) and np.array_equal(f(y), ref)

    dy, dym = pair((N, 2 * H, 2 * W, C))
    mask, maskm = pair((N, H, W, C))
    dx0, dx0m = pair((N, H, W, C))
    for acc in (0, 1):
        dx, dxm = dx0.clone(), dx0m.clone()
        S.upsample_bwd(dy, dx, mask, accumulate=bool(acc))
        S.upsample_bwd(dym, dxm, maskm, accumulate=bool(acc))
        torch.cuda.synchronize()
        d = f(dy).reshape(N, H, 2, W, 2, C)
        s4 = (((d[:, :, 0, :, 0] + d[:, :, 0, :, 1]) + d[:, :, 1, :, 0]) + d[:, :, 1, :, 1])
        s4 = np.where(f(mask) > 0, s4, 0.0) + (f(dx0) if acc else 0.0)
        ref = torch.from_numpy(s4.astype(np.float32)).to(torch.bfloat16).float().numpy()
        assert torch.equal(dx, dxm.contiguous()) and np.array_equal(f(dx), ref), acc

    src, srcm = pair((N, H, W, C))
    for acc in (0, 1):
        dst, dstm = dx0.clone(), dx0m.clone()
        S.masked_add(src, dst, mask, accumulate=bool(acc))
        S.masked_add(srcm, dstm, maskm, accumulate=bool(acc))
        torch.cuda.synchronize()
        v = np.where(f(mask) > 0, f(src), 0.0) + (f(dx0) if acc else 0.0)
        ref = torch.from_numpy(v.astype(np.float32)).to(torch.bfloat16).float().numpy()
        assert torch.equal(dst, dstm.contiguous()) and np.array_equal(f(dst), ref), acc
