"""Learned warm start for steady lid-driven cavities with obstacles.

A ~1.0 M-parameter convolutional encoder-decoder maps (solid mask, Reynolds
number) to the steady (rho, u) of a 256x256 D2Q9 lid cavity.  Training and
inference run entirely on the engine's hand-written BF16 MFMA kernels
(include/mfa_surrogate.h) with FP32 accumulation and an FP32 Adam master
copy; torch only owns the device memory.  The solver then starts from
f = feq(rho_pred, u_pred) and runs to the same steady-state criterion as a
cold start, so the plain LBM stays both the fallback and the arbiter of convergence.

Network (NHWC, 3x3 convolutions, ReLU except the head, additive skips):
  x0[8] -> e1[32] -> e1b[32] -s2-> d1[64] -> e2[64] -s2-> d2[128] -> e3[128]
  -s2-> d3[128] -> b1[128] -> b2[128] -> up+e3 -> c3[128] -> c3p[64]
  -> up+e2 -> c2[64] -> c2p[32] -> up+e1b -> c1[32] -> out[8]
Outputs 0..2 are (rho-1)/u_lid^2, u_x/u_lid, u_y/u_lid.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import sys
import time

import numpy as np

from . import abi

LAYERS = [  # name, cin, cout, stride, relu, input
    ("e1", 8, 16, 1, True, "x0"),         # 256^2
    ("e1b", 16, 16, 1, True, "e1"),
    ("d1", 16, 32, 2, True, "e1b"),       # 128^2
    ("e2", 32, 32, 1, True, "d1"),
    ("d2", 32, 64, 2, True, "e2"),        # 64^2
    ("e3", 64, 64, 1, True, "d2"),
    ("d3", 64, 96, 2, True, "e3"),        # 32^2
    ("e4", 96, 96, 1, True, "d3"),
    ("d4", 96, 120, 2, True, "e4"),       # 16^2
    ("e5", 120, 120, 1, True, "d4"),
    ("d5", 120, 120, 2, True, "e5"),      # 8^2: the bottleneck sees the whole cavity
    ("b1", 120, 120, 1, True, "d5"),
    ("b2", 120, 120, 1, True, "b1"),
    ("c5", 120, 96, 1, True, "u5"),       # 16^2
    ("c4", 96, 64, 1, True, "u4"),        # 32^2
    ("c3", 64, 32, 1, True, "u3"),        # 64^2
    ("c2", 32, 16, 1, True, "u2"),        # 128^2
    ("c1", 16, 16, 1, True, "u1"),        # 256^2
    ("out", 16, 8, 1, False, "c1"),
]
# additive skips: name -> (upsampled tensor, skip tensor)
SKIPS = {"u5": ("b2", "e5"), "u4": ("c5", "e4"), "u3": ("c4", "e3"), "u2": ("c3", "e2"),
         "u1": ("c2", "e1b")}
# scaled targets: (rho - 1) / (RHO_SCALE u_lid^2), u / u_lid
RHO_SCALE = 10.0
# tensors that are ReLU outputs (their gradient is masked by value > 0)
RELU_OUT = {name for name, _, _, _, relu, _ in LAYERS if relu}

_lib = None


class Segment(ctypes.Structure):
    """mfa_nn_segment: one tensor of the flat parameter buffers."""
    _fields_ = [("offset", ctypes.c_longlong), ("n", ctypes.c_longlong), ("cout", ctypes.c_int),
                ("cin", ctypes.c_int), ("weight_decay", ctypes.c_float), ("w_bf16", ctypes.c_void_p),
                ("wt_bf16", ctypes.c_void_p)]


def lib():
    global _lib
    if _lib is None:
        L = abi.lib()
        vp = ctypes.c_void_p
        i = ctypes.c_int
        f = ctypes.c_float
        ll = ctypes.c_longlong
        L.mfa_nn_error.restype = ctypes.c_char_p
        L.mfa_nn_conv3x3.argtypes = [vp, vp, vp, vp, vp, i, i, i, i, i, i, i, vp]
        L.mfa_nn_conv3x3_bwd_data.argtypes = [vp, vp, vp, vp, i, i, i, i, i, i, i, vp]
        L.mfa_nn_conv3x3_bwd_weight.argtypes = [vp, vp, vp, vp, i, i, i, i, i, i, vp]
        L.mfa_nn_fill.argtypes = [vp, ll, f, vp]
        L.mfa_nn_prep_weights.argtypes = [vp, vp, vp, i, i, vp]
        L.mfa_nn_upsample_add.argtypes = [vp, vp, vp, i, i, i, i, vp]
        L.mfa_nn_upsample_bwd.argtypes = [vp, vp, vp, i, i, i, i, i, vp]
        L.mfa_nn_masked_add.argtypes = [vp, vp, vp, ll, i, vp]
        L.mfa_nn_mse.argtypes = [vp, vp, vp, vp, vp, ll, i, i, f, vp]
        L.mfa_nn_adam.argtypes = [vp, vp, vp, vp, ll, f, f, f, f, i, f, vp]
        L.mfa_nn_make_input.argtypes = [vp, vp, vp, i, i, i, vp]
        L.mfa_nn_adam_multi.argtypes = [vp, vp, vp, vp, ctypes.POINTER(Segment), i, f, f, f, f, i, vp]
        for n in ("mfa_nn_conv3x3", "mfa_nn_conv3x3_bwd_data", "mfa_nn_conv3x3_bwd_weight",
                  "mfa_nn_fill", "mfa_nn_prep_weights", "mfa_nn_upsample_add",
                  "mfa_nn_upsample_bwd", "mfa_nn_masked_add", "mfa_nn_mse", "mfa_nn_adam",
                  "mfa_nn_make_input", "mfa_nn_adam_multi"):
            getattr(L, n).restype = ctypes.c_int
        _lib = L
    return _lib


def _ck(rc):
    if rc != 0:
        raise abi.MfaError(rc, lib().mfa_nn_error().decode())


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _s():
    return abi.stream_ptr()


# ---- thin kernel wrappers (NHWC tensors) ------------------------------------
def conv(x, w, b, y=None, yf=None, stride=1, relu=True):
    N, H, W, Cin = x.shape
    Cout = w.shape[0]
    _ck(lib().mfa_nn_conv3x3(_p(x), _p(w), _p(b), _p(y), _p(yf), N, H, W, Cin, Cout, stride,
                             1 if relu else 0, _s()))


def conv_bwd_data(dy, wt, dx, mask, stride, accumulate=False):
    N, H, W, Cin = dx.shape
    Cout = dy.shape[3]
    _ck(lib().mfa_nn_conv3x3_bwd_data(_p(dy), _p(wt), _p(dx), _p(mask), N, H, W, Cin, Cout, stride,
                                      1 if accumulate else 0, _s()))


def conv_bwd_weight(x, dy, dw, db, stride):
    N, H, W, Cin = x.shape
    Cout = dy.shape[3]
    _ck(lib().mfa_nn_conv3x3_bwd_weight(_p(x), _p(dy), _p(dw), _p(db), N, H, W, Cin, Cout, stride,
                                        _s()))


def fill(t, v=0.0):
    _ck(lib().mfa_nn_fill(_p(t), t.numel(), v, _s()))


def upsample_add(x, skip, y):
    N, H, W, C = x.shape
    _ck(lib().mfa_nn_upsample_add(_p(x), _p(skip), _p(y), N, H, W, C, _s()))


def upsample_bwd(dy, dx, mask, accumulate=False):
    N, H, W, C = dx.shape
    _ck(lib().mfa_nn_upsample_bwd(_p(dy), _p(dx), _p(mask), N, H, W, C, 1 if accumulate else 0,
                                  _s()))


def masked_add(src, dst, mask, accumulate=False):
    _ck(lib().mfa_nn_masked_add(_p(src), _p(dst), _p(mask), src.numel(), 1 if accumulate else 0,
                                _s()))


class WarmStartNet:
    """Parameters (FP32 master + BF16 working copies) and the hand-scheduled
    forward / backward passes over the MFMA kernels.

    The FP32 parameters, their gradients and Adam's two moments each live in
    one flat buffer, every layer's weight and bias a view into it, so zeroing
    the gradients is one launch and the optimiser step (which also writes the
    BF16 working copies) is another."""

    def __init__(self, H=256, W=256, seed=0, dev="cuda:0"):
        import torch
        self.torch = torch
        self.dev = torch.device(dev)
        self.H, self.W = H, W
        rng = np.random.default_rng(seed)
        total = sum(cout * 9 * cin + cout for _, cin, cout, _, _, _ in LAYERS)
        self.flat_w = torch.zeros(total, dtype=torch.float32, device=self.dev)
        self.flat_g = torch.zeros(total, dtype=torch.float32, device=self.dev)
        self.layers = {}
        off = 0
        for name, cin, cout, stride, relu, src in LAYERS:
            std = math.sqrt(2.0 / (9 * cin)) if relu else math.sqrt(1.0 / (9 * cin))
            w = rng.standard_normal((cout, 9 * cin)) * std
            if name == "e1":
                w.reshape(cout, 9, cin)[:, :, 5:] = 0.0     # padded input channels
            nw = cout * 9 * cin
            l = self.layers[name] = {
                "cin": cin, "cout": cout, "stride": stride, "relu": relu, "src": src,
                "off": off,
                "w32": self.flat_w[off:off + nw].view(cout, 9 * cin),
                "gw": self.flat_g[off:off + nw].view(cout, 9 * cin),
                "b32": self.flat_w[off + nw:off + nw + cout],
                "gb": self.flat_g[off + nw:off + nw + cout],
                "w": torch.empty((cout, 3, 3, cin), dtype=torch.bfloat16, device=self.dev),
                "wt": torch.empty((3, 3, cin, cout), dtype=torch.bfloat16, device=self.dev),
            }
            l["w32"].copy_(torch.tensor(w, dtype=torch.float32))
            off += nw + cout
        self.nparams = total
        self._adam = None
        self._segs = {}
        self.t = 0
        self.prep()

    # -- parameters --------------------------------------------------------
    def prep(self):
        for l in self.layers.values():
            _ck(lib().mfa_nn_prep_weights(_p(l["w32"]), _p(l["w"]), _p(l["wt"]), l["cout"], l["cin"],
                                          _s()))

    def state_dict(self):
        return {k: (l["w32"].cpu().numpy(), l["b32"].cpu().numpy()) for k, l in self.layers.items()}

    def load_state(self, sd):
        torch = self.torch
        for k, (w, b) in sd.items():
            self.layers[k]["w32"].copy_(torch.from_numpy(np.asarray(w, np.float32)))
            self.layers[k]["b32"].copy_(torch.from_numpy(np.asarray(b, np.float32)))
        self.prep()

    def save(self, path):
        sd = self.state_dict()
        np.savez_compressed(path, **{f"{k}.w": v[0] for k, v in sd.items()},
                            **{f"{k}.b": v[1] for k, v in sd.items()})

    def load(self, path):
        z = np.load(path)
        self.load_state({k: (z[f"{k}.w"], z[f"{k}.b"]) for k in self.layers})

    # -- forward -----------------------------------------------------------
    def shapes(self, N):
        H, W = self.H, self.W
        sh = {"x0": (N, H, W, 8)}
        for name, cin, cout, stride, relu, src in LAYERS:
            if src in SKIPS:
                a, b = SKIPS[src]
                sh[src] = sh[b][:3] + (sh[a][3],)
            n, h, w, _ = sh[src]
            sh[name] = (n, (h + stride - 1) // stride, (w + stride - 1) // stride, cout)
        return sh

    def forward(self, x0, keep=True):
        """x0: (N,H,W,8) bf16.  Returns the FP32 head output (N,H,W,8)."""
        torch = self.torch
        N = x0.shape[0]
        sh = self.shapes(N)
        act = {"x0": x0}
        for name, cin, cout, stride, relu, src in LAYERS:
            if src in SKIPS and src not in act:
                a, b = SKIPS[src]
                act[src] = torch.empty(sh[src], dtype=torch.bfloat16, device=self.dev)
                upsample_add(act[a], act[b], act[src])
            l = self.layers[name]
            if name == "out":
                yf = torch.empty(sh[name], dtype=torch.float32, device=self.dev)
                conv(act[src], l["w"], l["b32"], None, yf, stride, relu=False)
                act[name] = yf
            else:
                y = torch.empty(sh[name], dtype=torch.bfloat16, device=self.dev)
                conv(act[src], l["w"], l["b32"], y, None, stride, relu)
                act[name] = y
        self.act = act if keep else None
        return act["out"]

    # -- backward ----------------------------------------------------------
    def zero_grad(self):
        fill(self.flat_g)

    def backward(self, g_out):
        """g_out: (N,H,W,8) bf16 gradient of the loss w.r.t. the head output."""
        torch = self.torch
        act = self.act
        grad = {"out": g_out}
        # gradients are w.r.t. pre-activation of each layer's output
        for name, cin, cout, stride, relu, src in reversed(LAYERS):
            l = self.layers[name]
            g = grad.pop(name)
            conv_bwd_weight(act[src], g, l["gw"], l["gb"], stride)
            if name == "e1":
                break
            mask = act[src] if src in RELU_OUT else None
            acc = src in grad
            if not acc:
                grad[src] = torch.empty_like(act[src])
            conv_bwd_data(g, l["wt"], grad[src], mask, stride, accumulate=acc)
            if src in SKIPS:
                a, b = SKIPS[src]
                gu = grad.pop(src)
                ga = torch.empty_like(act[a])
                upsample_bwd(gu, ga, act[a])
                grad[a] = ga
                accb = b in grad
                if not accb:
                    grad[b] = torch.empty_like(act[b])
                masked_add(gu, grad[b], act[b], accumulate=accb)

    def _segments(self, wd):
        """The flat buffers' segment table for mfa_nn_adam_multi: per layer
        its weight (decayed by wd; BF16 copy and transpose written) and its
        bias (not decayed)."""
        if wd not in self._segs:
            segs = []
            for name, cin, cout, _, _, _ in LAYERS:
                l = self.layers[name]
                nw = cout * 9 * cin
                segs.append(Segment(l["off"], nw, cout, cin, wd, l["w"].data_ptr(), l["wt"].data_ptr()))
                segs.append(Segment(l["off"] + nw, cout, 0, 0, 0.0, None, None))
            self._segs[wd] = (Segment * len(segs))(*segs)
        return self._segs[wd]

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, wd=0.0):
        """One Adam step of every parameter (one launch), which also writes
        the BF16 working copies the convolutions read."""
        torch = self.torch
        if self._adam is None:
            self._adam = (torch.zeros_like(self.flat_w), torch.zeros_like(self.flat_w))
        self.t += 1
        segs = self._segments(float(wd))
        m, v = self._adam
        _ck(lib().mfa_nn_adam_multi(_p(self.flat_w), _p(self.flat_g), _p(m), _p(v), segs, len(segs),
                                    lr, beta1, beta2, eps, self.t, _s()))


# ---- data -------------------------------------------------------------------
def re_feature(Re):
    """log-Re mapped to [-1, 1] over [100, 3200]."""
    lo, hi = math.log(100.0), math.log(3200.0)
    return (2.0 * (np.log(np.asarray(Re, dtype=np.float64)) - lo) / (hi - lo) - 1.0).astype(np.float32)


def random_obstacles(rng, N=256, max_boxes=3, side=(8, 48)):
    m = np.zeros((N, N), dtype=np.uint8)
    for _ in range(int(rng.integers(0, max_boxes + 1))):
        w = int(rng.integers(side[0], side[1] + 1))
        h = int(rng.integers(side[0], side[1] + 1))
        x0 = int(rng.integers(0, N - w + 1))
        y0 = int(rng.integers(0, N - h + 1))
        m[y0:y0 + h, x0:x0 + w] = 1
    return m


def make_cases(n, seed, N=256, re_range=(100.0, 3200.0), u_lid=0.1):
    """Configuration 5 cases: obstacle masks, Re log-uniform, tau for lid u."""
    rng = np.random.default_rng(seed)
    masks = np.stack([random_obstacles(rng, N) for _ in range(n)])
    Re = np.exp(rng.uniform(math.log(re_range[0]), math.log(re_range[1]), size=n))
    tau = 0.5 + 3.0 * u_lid * N / Re
    return masks, Re, tau


def targets_from_fields(rho, vel, u_lid=0.1):
    """(n,H,W) rho, (n,2,H,W) u -> (n,H,W,3) float32 scaled targets."""
    t = np.stack([(rho - 1.0) / (RHO_SCALE * u_lid ** 2), vel[:, 0] / u_lid, vel[:, 1] / u_lid], axis=-1)
    return t.astype(np.float32)


def fields_from_prediction(pred, u_lid=0.1):
    """(n,H,W,>=3) scaled prediction -> rho (n,H,W), u (n,2,H,W) float64."""
    p = np.asarray(pred, dtype=np.float64)
    rho = 1.0 + p[..., 0] * (RHO_SCALE * u_lid ** 2)
    u = np.stack([p[..., 1], p[..., 2]], axis=1) * u_lid
    return rho, u


# ---- timing -----------------------------------------------------------------
def conv_flops(net, N):
    """FLOPs (2 per multiply-add) of one training iteration on a batch of N,
    per kernel family: forward of every layer, data gradient of every layer
    but the first (its input needs none), weight gradient of every layer."""
    sh = net.shapes(N)
    per = {name: 2 * 9 * cin * cout * sh[name][0] * sh[name][1] * sh[name][2]
           for name, cin, cout, stride, relu, src in LAYERS}
    return {"conv_fwd": sum(per.values()), "conv_bwd_data": sum(v for k, v in per.items() if k != "e1"),
            "conv_bwd_weight": sum(per.values())}


def bench(batch=16, iters=20, warmup=3, H=256, seed=0, dev="cuda:0"):
    """Training iterations (forward, masked MSE, both backward passes, Adam)
    of the warm-start network on random cavities: ms per iteration and
    aggregate TFLOP/s of the convolutions."""
    import json
    import time
    import torch
    net = WarmStartNet(H=H, W=H, seed=seed, dev=dev)
    rng = np.random.default_rng(seed)
    masks = torch.from_numpy(np.stack([random_obstacles(rng, H) for _ in range(batch)])).to(net.dev)
    re = torch.from_numpy(re_feature(np.exp(rng.uniform(math.log(100), math.log(3200), batch)))).to(net.dev)
    tgt = torch.from_numpy(rng.standard_normal((batch, H, H, 3)).astype(np.float32)).to(net.dev)
    x0 = torch.empty((batch, H, H, 8), dtype=torch.bfloat16, device=net.dev)
    g = torch.empty((batch, H, H, 8), dtype=torch.bfloat16, device=net.dev)
    part = torch.zeros((batch * H * H + 255) // 256, dtype=torch.float32, device=net.dev)
    _ck(lib().mfa_nn_make_input(_p(masks), _p(re), _p(x0), batch, H, H, _s()))
    count = float(batch * H * H * 3)

    def it():
        out = net.forward(x0)
        _ck(lib().mfa_nn_mse(_p(out), _p(tgt), _p(masks), _p(g), _p(part), batch * H * H, 8, 3,
                             1.0 / count, _s()))
        net.zero_grad()
        net.backward(g)
        net.adam_step(1e-4)
    for _ in range(warmup):
        it()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(iters):
        it()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / iters
    fl = conv_flops(net, batch)
    out = {"batch": batch, "H": H, "iters": iters, "ms_per_iteration": round(ms, 4),
           "conv_gflop_per_iteration": {k: round(v / 1e9, 2) for k, v in fl.items()},
           "conv_tflops_aggregate": round(sum(fl.values()) / (ms * 1e-3) / 1e12, 2),
           "wall_s": round(time.perf_counter() - t0, 3)}
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    import argparse
    p = argparse.ArgumentParser(description="time training iterations of the warm-start network")
    p.add_argument("--batch", type=int, default=16)
    p.add_argument("--iters", type=int, default=20)
    p.add_argument("--size", type=int, default=256)
    a = p.parse_args()
    bench(a.batch, a.iters, H=a.size)
