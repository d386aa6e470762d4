"""Learned warm start pipeline (configuration 5).

  generate : the batched engine runs N obstacle cavities (256^2, Re log-uniform
             in [100, 3200], lid u = 0.1) to a steady state and keeps (rho, u)
  train    : the BF16-MFMA encoder-decoder is fitted to those fields
  evaluate : held-out cases run to the 1e-8 criterion from rest (cold), from
             the prediction (feq, or feq plus its Chapman-Enskog
             non-equilibrium part) and from the cold start's own converged
             fields (the bound of a perfect prediction); steps, time to the
             criterion per case and per Re band, and the difference of the
             converged fields are reported

Usage: python -m muphfasa_amd.warmstart --stage train --n-train 4096 ...
       python -m muphfasa_amd.warmstart --stage eval --weights w.npz ...
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import abi
from . import ensemble as E
from . import surrogate as S


def _log(**kw):
    print(json.dumps(kw), flush=True)


def generate(n, seed, dtype="f32", tol=1e-6, max_steps=60000, check=1000, N=256, tile=(64, 4),
             max_seconds=None, cases=None):
    """Returns masks (n,N,N) uint8, Re (n,), fields on device (rho, vel) and
    the run statistics.  cases: (masks, Re, tau) instead of S.make_cases."""
    masks, Re, tau = cases if cases is not None else S.make_cases(n, seed, N)
    ens = E.Ensemble([(N, N)] * n, 1.0 / tau, dtype=dtype, tile=tile, masks=list(masks))
    ens.init()
    t0 = time.time()

    def log(s, nact, rmax):
        if s % (10 * check) == 0:
            _log(stage="generate", step=s, active=nact, max_res=rmax, t=round(time.time() - t0, 1))
    r = E.run_to_steady(ens, check, tol, max_steps, graph=True, log=log, max_seconds=max_seconds)
    ens.set_active(np.ones(n, np.int32))
    ens.moments()
    stats = {"n": n, "tol": tol, "max_steps": max_steps, "dtype": dtype,
             "converged": int((r["status"] == "converged").sum()),
             "diverged": int((r["status"] == "diverged").sum()),
             "median_steps": float(np.median(r["steps"])), "gpu_s": round(r["gpu_s"], 2),
             "unconverged_residual_p50_p90": [float(x) for x in np.percentile(
                 r["residual"][r["status"] == "max_steps"], [50, 90])]
             if (r["status"] == "max_steps").any() else None,
             "mlups": round(r["mlups_gpu"], 1)}
    return masks, Re, ens, r, stats


def device_targets(ens, n, N=256, u_lid=0.1):
    """(n,N,N,3) float32 scaled targets built on the device from the ensemble fields."""
    torch = ens.torch
    tc = ens.total_cells
    rho = ens.rho.view(n, N, N)
    ux = ens.vel[:tc].view(n, N, N)
    uy = ens.vel[tc:].view(n, N, N)
    t = torch.stack([(rho - 1.0) / (S.RHO_SCALE * u_lid ** 2), ux / u_lid, uy / u_lid], dim=-1)
    return t.to(torch.float32).contiguous()


def make_inputs(net, masks_dev, re_feat_dev, start, count):
    torch = net.torch
    N = net.H
    x0 = torch.empty((count, N, N, 8), dtype=torch.bfloat16, device=net.dev)
    S._ck(S.lib().mfa_nn_make_input(S._p(masks_dev[start:start + count]),
                                    S._p(re_feat_dev[start:start + count]), S._p(x0), count, N, N,
                                    S._s()))
    return x0


def flops_per_sample(net):
    """Forward FLOPs of one 256x256 sample (2 per multiply-add)."""
    sh = net.shapes(1)
    return sum(2 * 9 * cin * cout * sh[name][1] * sh[name][2]
               for name, cin, cout, stride, relu, src in S.LAYERS)


def train(net, masks_dev, re_feat_dev, targets, epochs=20, batch=16, lr=2e-3, seed=0, log_every=50,
          seconds=None):
    """Adam with a cosine schedule; stops after `epochs` or `seconds`."""
    torch = net.torch
    n = targets.shape[0]
    N = net.H
    rng = np.random.default_rng(seed)
    g = torch.empty((batch, N, N, 8), dtype=torch.bfloat16, device=net.dev)
    part = torch.zeros((batch * N * N + 255) // 256, dtype=torch.float32, device=net.dev)
    steps_per_epoch = n // batch
    total = epochs * steps_per_epoch
    hist = []
    t0 = time.time()
    it = 0
    x_buf = torch.empty((batch, N, N, 8), dtype=torch.bfloat16, device=net.dev)
    t_buf = torch.empty((batch, N, N, 3), dtype=torch.float32, device=net.dev)
    m_buf = torch.empty((batch, N, N), dtype=torch.uint8, device=net.dev)
    r_buf = torch.empty(batch, dtype=torch.float32, device=net.dev)
    fl = 3 * flops_per_sample(net) * batch          # forward + data grad + weight grad
    stop = False
    for ep in range(epochs):
        if stop:
            break
        perm = torch.from_numpy(rng.permutation(n)).to(net.dev)
        ep_loss = 0.0
        for bi in range(steps_per_epoch):
            idx = perm[bi * batch:(bi + 1) * batch]
            torch.index_select(masks_dev, 0, idx, out=m_buf)
            torch.index_select(re_feat_dev, 0, idx, out=r_buf)
            torch.index_select(targets, 0, idx, out=t_buf)
            S._ck(S.lib().mfa_nn_make_input(S._p(m_buf), S._p(r_buf), S._p(x_buf), batch, N, N,
                                            S._s()))
            frac = it / total if seconds is None else min(1.0, (time.time() - t0) / seconds)
            lr_t = lr * 0.5 * (1 + math.cos(math.pi * frac))
            out = net.forward(x_buf)
            count = float(batch * N * N * 3)
            S._ck(S.lib().mfa_nn_mse(S._p(out), S._p(t_buf), S._p(m_buf), S._p(g), S._p(part),
                                     batch * N * N, 8, 3, 1.0 / count, S._s()))
            net.zero_grad()
            net.backward(g)
            net.adam_step(lr_t)
            if it % log_every == 0 or bi == steps_per_epoch - 1:
                l = part.sum().item() / count
                ep_loss = l
                hist.append((it, l))
                if it % (4 * log_every) == 0:
                    _log(stage="train", epoch=ep, it=it, loss=l, lr=lr_t, t=round(time.time() - t0, 1))
            it += 1
            if seconds is not None and time.time() - t0 > seconds:
                stop = True
                break
        torch.cuda.synchronize()
        el = time.time() - t0
        _log(stage="train", epoch=ep, loss=ep_loss, t=round(el, 1),
             tflops=round(fl * it / el / 1e12, 2))
    return hist, time.time() - t0


def predict(net, masks_dev, re_feat_dev, batch=16):
    """Scaled predictions (n,N,N,8) float32 on the device."""
    torch = net.torch
    n = masks_dev.shape[0]
    N = net.H
    out = torch.empty((n, N, N, 8), dtype=torch.float32, device=net.dev)
    for s in range(0, n, batch):
        c = min(batch, n - s)
        x0 = make_inputs(net, masks_dev, re_feat_dev, s, c)
        out[s:s + c] = net.forward(x0, keep=False)
    return out


def fields_on_device(pred, masks_dev, u_lid=0.1):
    """rho (n*N*N,), vel (2*n*N*N,) float64 in the ensemble's cell order.

    The predicted density is shifted so that each cavity's fluid holds the
    mass of the cold start (rho = 1 per fluid cell): a closed cavity keeps its
    initial mass, and its steady velocity scales with it, so a start with any
    other mass would converge to a different field."""
    import torch
    p = pred.double()
    fluid = (masks_dev == 0).double()
    drho = p[..., 0] * (S.RHO_SCALE * u_lid ** 2) * fluid
    nfl = fluid.sum(dim=(1, 2)).clamp_min(1.0)
    drho = drho - (drho.sum(dim=(1, 2)) / nfl)[:, None, None] * fluid
    rho = 1.0 + drho
    ux = p[..., 1] * u_lid * fluid
    uy = p[..., 2] * u_lid * fluid
    return rho.reshape(-1).contiguous(), torch.cat([ux.reshape(-1), uy.reshape(-1)]).contiguous()


MODES = ("cold", "warm", "warm_neq", "warm_plain", "exact", "exact_neq")
RE_BANDS = (100.0, 200.0, 400.0, 800.0, 1600.0, 3200.0)


def _start_fields(mode, net, masks_dev, re_dev, cold_fields):
    """(rho, vel) on the device for a start mode, and the inference time."""
    import torch
    if mode.startswith("exact"):
        rho, vel = cold_fields
        return rho.clone(), vel.clone(), 0.0
    ti = time.perf_counter()
    pred = predict(net, masks_dev, re_dev)
    rho, vel = fields_on_device(pred, masks_dev)
    torch.cuda.synchronize()
    return rho, vel, time.perf_counter() - ti


def evaluate(net, n_eval=256, seed=777, tol=1e-8, max_steps=200000, check=1000, dtype="f64", N=256,
             tile=(64, 4), max_seconds=None, modes=MODES, cases=None):
    """Runs the held-out cases to the steady criterion from each start:
      cold       rest state (rho = 1, u = 0), the plain method
      warm_plain feq of the network's prediction (mass restored)
      warm       the same followed by one step and the mean of the two
                 consecutive states (mfa_ensemble_init_warm)
      warm_neq   feq plus the first-order Chapman-Enskog non-equilibrium part
                 of the prediction, then the two-state average
                 (mfa_ensemble_init_start, MFA_START_NEQ | MFA_START_AVERAGE)
      exact, exact_neq
                 the warm and warm_neq recipes applied to the cold start's own
                 converged fields: what a perfect prediction would gain
    and compares steps, time to the criterion and converged fields with the
    cold start.  Per case, the time to the criterion is its steps times its
    share of the batch's measured step rate (cells / lattice updates per
    second) plus its share of the network inference."""
    import torch
    masks, Re, tau = cases if cases is not None else S.make_cases(n_eval, seed, N)
    masks_dev = torch.from_numpy(masks).to(net.dev)
    re_dev = torch.from_numpy(S.re_feature(Re)).to(net.dev)
    modes = tuple(modes)
    if any(m.startswith("exact") for m in modes) and "cold" not in modes:
        raise ValueError("the exact starts need the cold run")
    # untimed warm-up of the ensemble kernels (a fresh device runs its first
    # few thousand steps slower)
    ew = E.Ensemble([(N, N)] * min(n_eval, 32), 1.0 / tau[:min(n_eval, 32)], dtype=dtype, tile=tile,
                    masks=list(masks[:min(n_eval, 32)]))
    ew.init()
    ew.run(2000, graph=True)
    torch.cuda.synchronize()
    del ew
    if any(m.startswith("warm") for m in modes):
        # and of the network kernels and the field conversion after them, so
        # that the timed inference is steady
        fields_on_device(predict(net, masks_dev[:16], re_dev[:16]), masks_dev[:16])
        torch.cuda.synchronize()
    res = {}
    vel_pred = None
    cold_fields = None
    order = (["cold"] if "cold" in modes else []) + [m for m in modes if m != "cold"]
    for mode in order:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ens = E.Ensemble([(N, N)] * n_eval, 1.0 / tau, dtype=dtype, tile=tile, masks=list(masks))
        t_inf = 0.0
        if mode != "cold":
            rho, vel, t_inf = _start_fields(mode, net, masks_dev, re_dev, cold_fields)
            ens.init(rho, vel, average=(mode != "warm_plain"), neq=mode.endswith("_neq"))
            if mode.startswith("warm"):
                vel_pred = vel
            del rho
        else:
            ens.init()

        def log(s, nact, rmax, mode=mode):
            if s % (20 * check) == 0:
                _log(stage="eval", mode=mode, step=s, active=nact, max_res=rmax,
                     t=round(time.perf_counter() - t0, 1))
        r = E.run_to_steady(ens, check, tol, max_steps, graph=True, log=log, max_seconds=max_seconds)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ens.set_active(np.ones(n_eval, np.int32))
        ens.moments()
        rate = r["lattice_updates"] / max(r["gpu_s"], 1e-9)        # lattice updates per second
        t_case = r["steps"] * float(N * N) / rate + t_inf / n_eval
        res[mode] = {"steps": r["steps"], "status": r["status"], "wall_s": wall, "gpu_s": r["gpu_s"],
                     "inference_s": t_inf, "vel": ens.vel.clone(), "t_case": t_case,
                     "lattice_updates": r["lattice_updates"]}
        if mode == "cold":
            cold_fields = (ens.rho.clone(), ens.vel.clone())
        _log(stage="eval_mode_done", mode=mode, wall_s=round(wall, 2),
             converged=int((r["status"] == "converged").sum()),
             steps_total=int(r["steps"].sum()), glups=round(rate / 1e9, 2))
        del ens
    tc = n_eval * N * N
    cc = res.get("cold")
    summary = {"n_eval": n_eval, "tol": tol, "max_steps": max_steps, "dtype": dtype,
               "check": check}
    per_case = {"Re": Re, "solid_cells": masks.reshape(n_eval, -1).sum(axis=1)}
    for mode, rm in res.items():
        summary[f"{mode}_converged"] = int((rm["status"] == "converged").sum())
        summary[f"{mode}_steps_total"] = int(rm["steps"].sum())
        summary[f"{mode}_wall_s"] = round(rm["wall_s"], 2)
        summary[f"{mode}_lattice_updates"] = int(rm["lattice_updates"])
        per_case[f"steps_{mode}"] = rm["steps"]
        per_case[f"status_{mode}"] = rm["status"].astype(str)
        per_case[f"time_{mode}"] = rm["t_case"]
    if cc is not None:
        uc = cc["vel"].view(2, n_eval, -1)
        if vel_pred is not None:
            # surrogate error against the converged cold field, per case
            dp = (vel_pred - cc["vel"]).view(2, n_eval, -1)
            per_case["pred_err_max"] = dp.abs().amax(dim=(0, 2)).cpu().numpy() / 0.1
            per_case["pred_err_l2"] = torch.sqrt(
                (dp * dp).sum(dim=(0, 2)) / (uc * uc).sum(dim=(0, 2)).clamp_min(1e-300)).cpu().numpy()
        bands = {}
        for mode, rm in res.items():
            if mode == "cold":
                continue
            both = (rm["status"] == "converged") & (cc["status"] == "converged")
            dv = (rm["vel"] - cc["vel"]).abs()
            dmax = torch.maximum(dv[:tc], dv[tc:]).view(n_eval, -1).max(dim=1).values.cpu().numpy() / 0.1
            per_case[f"field_diff_max_{mode}"] = dmax
            ratio = cc["steps"][both] / np.maximum(rm["steps"][both], 1)
            tratio = cc["t_case"][both] / np.maximum(rm["t_case"][both], 1e-12)
            summary[f"{mode}_both_converged"] = int(both.sum())
            summary[f"{mode}_only_converged"] = int(((rm["status"] == "converged") &
                                                     (cc["status"] != "converged")).sum())
            summary[f"{mode}_steps_ratio_median"] = float(np.median(ratio)) if both.any() else None
            summary[f"{mode}_steps_ratio_mean"] = float(np.mean(ratio)) if both.any() else None
            summary[f"{mode}_steps_ratio_of_sums"] = (
                float(cc["steps"][both].sum() / max(rm["steps"][both].sum(), 1)) if both.any() else None)
            summary[f"{mode}_time_ratio_median"] = float(np.median(tratio)) if both.any() else None
            summary[f"{mode}_time_ratio_ge3_frac"] = float(np.mean(tratio >= 3.0)) if both.any() else None
            summary[f"{mode}_wall_ratio"] = round(cc["wall_s"] / rm["wall_s"], 3)
            summary[f"{mode}_lu_ratio"] = round(cc["lattice_updates"] / max(rm["lattice_updates"], 1), 3)
            summary[f"{mode}_inference_s"] = round(rm["inference_s"], 3)
            summary[f"{mode}_field_diff_max"] = float(dmax[both].max()) if both.any() else None
            summary[f"{mode}_field_diff_median"] = float(np.median(dmax[both])) if both.any() else None
            summary[f"{mode}_field_diff_p90"] = float(np.percentile(dmax[both], 90)) if both.any() else None
            rows = []
            for lo, hi in zip(RE_BANDS[:-1], RE_BANDS[1:]):
                inb = (Re >= lo) & (Re < hi)
                sel = both & inb
                rows.append({
                    "Re": [lo, hi], "cases": int(inb.sum()), "both_converged": int(sel.sum()),
                    "cold_capped": int((inb & (cc["status"] != "converged")).sum()),
                    "start_capped": int((inb & (rm["status"] != "converged")).sum()),
                    "time_ratio_median": float(np.median(cc["t_case"][sel] / rm["t_case"][sel]))
                    if sel.any() else None,
                    "steps_ratio_median": float(np.median(cc["steps"][sel] / np.maximum(rm["steps"][sel], 1)))
                    if sel.any() else None,
                    "cold_time_median_s": float(np.median(cc["t_case"][sel])) if sel.any() else None,
                    "start_time_median_s": float(np.median(rm["t_case"][sel])) if sel.any() else None,
                    "pred_err_l2_median": float(np.nanmedian(per_case["pred_err_l2"][inb]))
                    if "pred_err_l2" in per_case and inb.any() else None,
                })
            bands[mode] = rows
        summary["bands"] = bands
        if "pred_err_l2" in per_case:
            # nan-aware: a lattice that blew up has no field to compare with
            summary["pred_err_l2_median"] = float(np.nanmedian(per_case["pred_err_l2"]))
            summary["pred_err_max_median"] = float(np.nanmedian(per_case["pred_err_max"]))
    return summary, per_case


def target_error(n, seed, gen_tol=2e-6, gen_max_steps=200000, ref_tol=1e-8, N=256):
    """How far the training targets are from the steady state: n cavities
    (the training distribution, another seed) generated as for training (FP32
    storage to gen_tol) and, as the reference, in FP64 to ref_tol; per case
    max over fluid cells of |u - u_ref| / U and the relative L2 distance of
    u, over the cases whose reference converged, by Re band."""
    import torch
    cases = S.make_cases(n, seed, N)
    fields = {}
    runs = {}
    for dt, tol in (("f64", ref_tol), ("f32", gen_tol)):
        masks, Re, ens, r, stats = generate(n, seed, dtype=dt, tol=tol, max_steps=gen_max_steps, N=N,
                                            cases=cases)
        _log(stage="target_error_run", **stats)
        fields[dt] = ens.vel.clone()
        runs[dt] = r
        del ens
        torch.cuda.empty_cache()
    masks, Re, tau = cases
    tc = n * N * N
    d = (fields["f32"] - fields["f64"]).abs() / 0.1
    dmax = torch.maximum(d[:tc], d[tc:]).view(n, -1).max(dim=1).values.cpu().numpy()
    dv = (fields["f32"] - fields["f64"]).view(2, n, -1)
    uv = fields["f64"].view(2, n, -1)
    l2 = torch.sqrt((dv * dv).sum(dim=(0, 2)) / (uv * uv).sum(dim=(0, 2)).clamp_min(1e-300)).cpu().numpy()
    ok = (runs["f64"]["status"] == "converged") & np.isfinite(dmax)
    bands = []
    for lo, hi in zip(RE_BANDS[:-1], RE_BANDS[1:]):
        sel = ok & (Re >= lo) & (Re < hi)
        if sel.any():
            bands.append({"Re": [lo, hi], "cases": int(sel.sum()),
                          "max_du_over_U_median": float(np.median(dmax[sel])),
                          "rel_l2_median": float(np.median(l2[sel]))})
    out = {"cases": n, "reference_converged": int(ok.sum()),
           "f32_converged": int((runs["f32"]["status"] == "converged").sum()),
           "max_du_over_U": {"median": float(np.median(dmax[ok])), "p90": float(np.percentile(dmax[ok], 90)),
                             "max": float(dmax[ok].max())} if ok.any() else None,
           "rel_l2_u": {"median": float(np.median(l2[ok])), "max": float(l2[ok].max())} if ok.any() else None,
           "bands": bands}
    _log(stage="target_error_done", **out)
    return out


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--stage", default="all", choices=["train", "eval", "all", "target-error"])
    p.add_argument("--n-target", type=int, default=64, help="cases of --stage target-error")
    p.add_argument("--n-train", type=int, default=4096)
    p.add_argument("--gen-tol", type=float, default=1e-6)
    p.add_argument("--gen-max-steps", type=int, default=60000)
    p.add_argument("--gen-seconds", type=float, default=None, help="wall budget of the generation")
    p.add_argument("--keep-residual", type=float, default=1e-3,
                   help="train only on cases whose last residual is below this")
    p.add_argument("--epochs", type=int, default=20)
    p.add_argument("--batch", type=int, default=16)
    p.add_argument("--lr", type=float, default=2e-3)
    p.add_argument("--train-seconds", type=float, default=None)
    p.add_argument("--weights", default="warmstart_weights.npz")
    p.add_argument("--n-eval", type=int, default=256)
    p.add_argument("--eval-max-steps", type=int, default=200000)
    p.add_argument("--tol", type=float, default=1e-8)
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--eval-out", default="", help="per-case results (.npz)")
    p.add_argument("--eval-seconds", type=float, default=None, help="wall budget per start mode")
    p.add_argument("--modes", default=",".join(MODES),
                   help="starts to evaluate (comma list of " + ", ".join(MODES) + ")")
    a = p.parse_args()
    import torch
    if a.stage == "target-error":
        target_error(a.n_target, a.seed + 7, gen_tol=a.gen_tol, gen_max_steps=a.gen_max_steps, ref_tol=a.tol)
        return
    net = S.WarmStartNet(seed=a.seed)
    _log(stage="init", params=net.nparams)
    if a.stage in ("train", "all"):
        masks, Re, ens, r, stats = generate(a.n_train, a.seed, tol=a.gen_tol, max_steps=a.gen_max_steps,
                                            max_seconds=a.gen_seconds)
        _log(stage="generate_done", **stats)
        targets = device_targets(ens, a.n_train)
        del ens
        torch.cuda.empty_cache()
        # a lattice that blew up (BGK at tau near 1/2 with a one-cell gap under
        # the lid, for instance) has no steady state to learn: drop it
        keep = torch.isfinite(targets).view(a.n_train, -1).all(dim=1).cpu().numpy()
        keep &= r["status"] != "diverged"
        # a flow still changing by more than keep_residual per check interval
        # (an unsteady one, behind an obstacle at high Re) is no steady state
        keep &= r["residual"] < a.keep_residual
        if not keep.all():
            _log(stage="generate_dropped", count=int((~keep).sum()),
                 cases=[{"i": int(i), "Re": float(Re[i])} for i in np.nonzero(~keep)[0][:20]])
            idx = torch.from_numpy(np.nonzero(keep)[0]).to(net.dev)
            targets = targets.index_select(0, idx).contiguous()
            masks, Re = masks[keep], Re[keep]
        masks_dev = torch.from_numpy(masks).to(net.dev)
        re_dev = torch.from_numpy(S.re_feature(Re)).to(net.dev)
        hist, t_train = train(net, masks_dev, re_dev, targets, a.epochs, a.batch, a.lr, a.seed,
                              seconds=a.train_seconds)
        os.makedirs(os.path.dirname(a.weights) or ".", exist_ok=True)
        net.save(a.weights)
        _log(stage="train_done", seconds=round(t_train, 1), final_loss=hist[-1][1],
             first_loss=hist[0][1], weights=a.weights)
        del targets
        torch.cuda.empty_cache()
    if a.stage in ("eval", "all"):
        if a.stage == "eval" and any(m.startswith("warm") for m in a.modes.split(",")):
            net.load(a.weights)
        summary, per_case = evaluate(net, a.n_eval, seed=a.seed + 1, tol=a.tol,
                                     max_steps=a.eval_max_steps, max_seconds=a.eval_seconds,
                                     modes=tuple(a.modes.split(",")))
        _log(stage="eval_done", **summary)
        if a.eval_out:
            np.savez(a.eval_out, **per_case)


if __name__ == "__main__":
    main()
