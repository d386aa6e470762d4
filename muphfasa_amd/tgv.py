"""3-D Taylor-Green vortex precision study (configuration 4).

D3Q27 BGK on a periodic N^3 box (default 384^3), u0 = 0.05, Re = 800 on
L = N / (2 pi) (nu = u0 L / Re, tau = 0.5 + 3 nu), run for the given number
of steps with populations stored as FP64 (the reference arithmetic), FP32,
FP16 and BF16 shifted deviations (collided in FP32).  Reports MLUPS per
storage type and, against the FP64 run, the relative L2 error of u and the
kinetic-energy history.
"""
from __future__ import annotations

import argparse
import json
import math
import time

import numpy as np

from . import abi

TORCH_DT = {"f64": "float64", "f32": "float32", "f16dev": "float16", "bf16dev": "bfloat16"}


def tgv_fields(torch, N, u0, dev):
    x = (torch.arange(N, dtype=torch.float64, device=dev) + 0.5) * (2 * math.pi / N)
    Z, Y, X = torch.meshgrid(x, x, x, indexing="ij")
    vel = torch.stack([u0 * torch.sin(X) * torch.cos(Y) * torch.cos(Z),
                       -u0 * torch.cos(X) * torch.sin(Y) * torch.cos(Z),
                       torch.zeros_like(X)]).contiguous()
    rho = (1.0 + 3.0 * u0 ** 2 / 16.0 * (torch.cos(2 * X) + torch.cos(2 * Y)) *
           (torch.cos(2 * Z) + 2.0))[None].contiguous()
    return rho, vel


def _pairs_offered(fa, fb, g, omega):
    """Whether mfa_step2 (two steps per launch) takes this lattice and
    storage; it returns MFA_ERR_UNSUPPORTED where it would run slower."""
    try:
        abi.step2(fa, fb, g, omega)
        return True
    except abi.MfaError as e:
        if e.code != -2:
            raise
        return False


def run(N=384, u0=0.05, Re=800.0, steps=5000, sample=500, storages=("f64", "f32", "f16dev", "bf16dev"),
        dev="cuda:0", layout="row", tune=False, pairs=True):
    """tune: time every launch candidate for each storage type on this box
    first (muphfasa_amd.autotune) and run with the fastest.  pairs: take the
    steps two per launch (mfa_step2) where the library offers it."""
    import torch
    L = N / (2 * math.pi)
    nu = u0 * L / Re
    tau = 0.5 + 3.0 * nu
    omega = 1.0 / tau
    n = [N, N, N]
    g = abi.resolve_bc(3, [abi.BC["Periodic"]] * 6)
    results = {}
    ref_vel = {}
    tuned = {}
    for st in storages:
        if tune:
            from . import autotune
            if st == "f64":
                best = autotune.tune_step("D3Q27", n, forced=False, layout=layout)[0]
            else:
                best = autotune.tune_step_dev("D3Q27", st, n, layout=layout)[0]
            tuned[st] = best
            print(json.dumps({"storage": st, "tuned": best}), flush=True)
            torch.cuda.empty_cache()
        tdt = getattr(torch, TORCH_DT[st])
        shape = abi.row_shape(27, n, 0, 3) if layout == "row" else (27, N, N, N)
        a = torch.empty(shape, dtype=tdt, device=dev)
        b = torch.empty_like(a)
        fa = abi.field(a, n, 0, 3, abi.DTYPE[st], layout=layout)
        fb = abi.field(b, n, 0, 3, abi.DTYPE[st], layout=layout)
        rho, vel = tgv_fields(torch, N, u0, dev)
        fr, fv = abi.field(rho, n, 0, 3), abi.field(vel, n, 0, 3)
        use_pairs = pairs and _pairs_offered(fa, fb, g, omega)
        abi.init_equilibrium(fr, fv, fa, 3)
        out = torch.zeros(8, dtype=torch.float64, device=dev)
        abi.reduce(fa, 3, abi.RED["mass"], out.data_ptr())
        torch.cuda.synchronize()
        m0 = out[0].item()
        ke, errs, t_kernel = [], [], 0.0
        done = 0
        while done < steps:
            k = min(sample, steps - done)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            if use_pairs:
                for _ in range(k // 2):
                    abi.step2(fa, fb, g, omega)
                    fa, fb = fb, fa
            for _ in range(k % 2 if use_pairs else k):
                abi.step(fa, fb, g, omega)
                fa, fb = fb, fa
            e1.record()
            torch.cuda.synchronize()
            t_kernel += e0.elapsed_time(e1) * 1e-3
            done += k
            abi.moments(fa, fr, fv, g)
            energy = 0.5 * (rho[0] * (vel * vel).sum(0)).mean().item() / (u0 ** 2)
            ke.append((done, energy))
            if st == "f64":
                ref_vel[done] = vel.clone()
            elif done in ref_vel:
                d = vel - ref_vel[done]
                errs.append((done, math.sqrt((d * d).sum().item() / (ref_vel[done] ** 2).sum().item())))
        abi.reduce(fa, 3, abi.RED["mass"], out.data_ptr())
        torch.cuda.synchronize()
        m1 = out[0].item()
        bytes_lu = 2 * 27 * a.element_size()
        mlups = N ** 3 * steps / t_kernel / 1e6
        launch = abi.get_launch(3) if st == "f64" else abi.get_launch_dev(3, abi.DTYPE[st])
        results[st] = {"mlups": round(mlups, 1), "bytes_per_lu": bytes_lu, "launch": list(launch),
                       "steps_per_launch": 2 if use_pairs else 1,
                       "GBs": round(mlups * bytes_lu / 1e3, 1), "ke": ke,
                       "l2_u_err_vs_f64": errs, "mass_rel_drift": abs(m1 - m0) / m0,
                       "finite": bool(torch.isfinite(vel).all().item())}
        if st != "f64":
            ke64 = dict(results["f64"]["ke"])
            results[st]["ke_rel_err_final"] = abs(ke[-1][1] - ke64[ke[-1][0]]) / ke64[ke[-1][0]]
        print(json.dumps({"storage": st, **{k: v for k, v in results[st].items()
                                             if k not in ("ke", "l2_u_err_vs_f64")},
                          "l2_u_err_final": errs[-1][1] if errs else 0.0,
                          "ke_final": ke[-1][1]}), flush=True)
        del a, b
        torch.cuda.empty_cache()
    return {"N": N, "u0": u0, "Re": Re, "tau": tau, "steps": steps, "layout": layout,
            "results": results}


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--n", type=int, default=384)
    p.add_argument("--steps", type=int, default=5000)
    p.add_argument("--sample", type=int, default=500)
    p.add_argument("--out", default="")
    p.add_argument("--layout", default="row", choices=["row", "component"])
    p.add_argument("--storages", default="f64,f32,f16dev,bf16dev")
    p.add_argument("--tune", action="store_true", help="autotune the launch per storage first")
    p.add_argument("--kernel", default="pair", choices=["pair", "single"],
                   help="pair: two steps per launch where mfa_step2 is offered")
    a = p.parse_args()
    r = run(a.n, steps=a.steps, sample=a.sample, storages=tuple(a.storages.split(",")),
            layout=a.layout, tune=a.tune, pairs=a.kernel == "pair")
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(r, fh)
    print(json.dumps({"tau": r["tau"], "summary": {k: {"mlups": v["mlups"],
                      "l2_u_err_final": v["l2_u_err_vs_f64"][-1][1] if v["l2_u_err_vs_f64"] else 0.0,
                      "ke_final": v["ke"][-1][1]} for k, v in r["results"].items()}}), flush=True)


if __name__ == "__main__":
    main()
