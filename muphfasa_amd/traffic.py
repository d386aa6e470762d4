"""Short runs of the engine's workloads for the two counter passes
(rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE) that measure HBM traffic per
launch.  Each run first streams the plain copy kernel over a known number of
bytes (the calibration muphfasa_amd.pmc divides by), then takes a few
launches of the workload's step kernel:

  ensemble512  512 D2Q9 lid cavities of the batched workload, FP64,
               two steps per launch (k_batch_step2)
  tgv27        D3Q27 Taylor-Green vortex 384^3, FP64, two steps per launch
               (k_step2p)
  tgv27_f32    the same stored as FP32 deviations (k_step2p_dev)
  tgv27_f16    the same stored as FP16 deviations, one step per launch
               (k_step_dev)
  tgv27_bf16   the same stored as BF16 deviations, one step per launch
  level64      D3Q19 256^3 FP64 forced, 64 boxes of 64^3 (k_level_step and
               the ghost exchange)

It prints one JSON line with what muphfasa_amd.pmc needs to turn the counters
into bytes per launch (kernel name fragment, cells, Q, bytes per value, steps
per launch, calibration bytes).

    rocprofv3 --pmc FETCH_SIZE -d OUT -o run --output-format csv -- \\
        python3 -m muphfasa_amd.traffic ensemble512
"""
from __future__ import annotations

import argparse
import json
import math

from . import abi


def _copy_probe(nbytes):
    from . import autotune
    return autotune.copy_ceiling(nbytes, reps=3)


def ensemble512(launches):
    import torch
    from . import ensemble as E
    N, Re, tau = E.cavity_ensemble_params(512)
    sizes = [(int(x), int(x)) for x in N]
    cells = sum(a * b for a, b in sizes)
    _copy_probe(cells * 9 * 8)
    ens = E.Ensemble(sizes, 1.0 / tau, dtype="f64")
    ens.init()
    for _ in range(launches):
        ens.step(pair=True)
    torch.cuda.synchronize()
    return {"kernel": "k_batch_step2", "cells": cells, "q": 9, "elem": 8, "steps_per_launch": 2,
            "copy_bytes": cells * 9 * 8, "workload": "512 D2Q9 lid cavities (side 64-256), FP64"}


def _tgv(st, launches, N=384):
    import torch
    from . import tgv
    n = [N, N, N]
    tdt = getattr(torch, tgv.TORCH_DT[st])
    elem = {"f64": 8, "f32": 4}.get(st, 2)
    pair = st in ("f64", "f32")             # 16-bit words take single steps
    _copy_probe(N ** 3 * 27 * elem)
    a = torch.empty(abi.row_shape(27, n, 0, 3), dtype=tdt, device="cuda:0")
    b = torch.empty_like(a)
    fa = abi.field(a, n, 0, 3, abi.DTYPE[st], layout="row")
    fb = abi.field(b, n, 0, 3, abi.DTYPE[st], layout="row")
    rho, vel = tgv.tgv_fields(torch, N, 0.05, "cuda:0")
    abi.init_equilibrium(abi.field(rho, n, 0, 3), abi.field(vel, n, 0, 3), fa, 3)
    del rho, vel
    L = N / (2 * math.pi)
    omega = 1.0 / (0.5 + 3.0 * 0.05 * L / 800.0)
    g = abi.resolve_bc(3, [abi.BC["Periodic"]] * 6)
    for _ in range(launches):
        (abi.step2 if pair else abi.step)(fa, fb, g, omega)
        fa, fb = fb, fa
    torch.cuda.synchronize()
    names = {"f64": "FP64", "f32": "FP32 deviations", "f16dev": "FP16 deviations",
             "bf16dev": "BF16 deviations"}
    return {"kernel": "k_step2p" if st == "f64" else ("k_step2p_dev" if pair else "k_step_dev"),
            "cells": N ** 3, "q": 27, "elem": elem, "steps_per_launch": 2 if pair else 1,
            "copy_bytes": N ** 3 * 27 * elem, "workload": f"D3Q27 TGV {N}^3, Re 800, {names[st]}"}


def level64(launches):
    import torch
    from . import level
    n = 256
    _copy_probe(n ** 3 * 19 * 8)
    level.run(n, 64, steps=2 * launches)
    torch.cuda.synchronize()
    return {"kernel": "k_level_step", "cells": n ** 3, "q": 19, "elem": 8, "steps_per_launch": 1,
            "copy_bytes": n ** 3 * 19 * 8,
            "workload": "D3Q19 256^3 FP64 forced periodic, 64 boxes of 64^3 (component-major)"}


WORKLOADS = {"ensemble512": ensemble512, "tgv27": lambda k: _tgv("f64", k),
             "tgv27_f32": lambda k: _tgv("f32", k), "tgv27_f16": lambda k: _tgv("f16dev", k),
             "tgv27_bf16": lambda k: _tgv("bf16dev", k), "level64": level64}


def main():
    p = argparse.ArgumentParser()
    p.add_argument("workload", choices=sorted(WORKLOADS))
    p.add_argument("--launches", type=int, default=8)
    a = p.parse_args()
    print(json.dumps(WORKLOADS[a.workload](a.launches)), flush=True)


if __name__ == "__main__":
    main()
