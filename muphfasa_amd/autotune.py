"""Autotuner for the fused step kernel's launch configuration.

Times mfa_step on device-resident populations of the given lattice and box
for every candidate configuration (block shape, cells per thread along z,
non-temporal loads) and returns them sorted by time.  ``apply=True`` installs
the fastest through mfa_set_launch.  The same routine sizes the batched
engine's tiles (muphfasa_amd.ensemble).
"""
from __future__ import annotations

import itertools
import json
import sys

from . import abi

CANDIDATES_3D = [(bx, by, nc, nt) for bx, by in ((256, 1), (128, 2), (64, 4), (64, 2), (128, 1))
                 for nc in (1, 2) for nt in (0, 1, 2)]
CANDIDATES_2D = [(bx, by, 1, nt) for bx, by in ((256, 1), (128, 2), (64, 4), (64, 2), (64, 1))
                 for nt in (0, 1, 2)]


def tune_step(lattice: str, n, forced: bool = True, periodic: bool = True, steps: int = 20,
              candidates=None, apply: bool = True, dev="cuda:0", layout: str = "row"):
    """layout: population order of the timed fields ("row" = the engine's
    row-interleaved allocation, "component" = the MultiFab's planes)."""
    import torch
    lid = abi.LATTICE[lattice]
    D, Q = abi.LATTICE_DQ[lid]
    n = list(n) + [1] * (3 - len(n))
    shape = abi.row_shape(Q, n, 0, D) if layout == "row" else (Q, n[2], n[1], n[0])
    a = torch.empty(shape, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    a.fill_(1.0 / Q)
    fa, fb = abi.field(a, n, 0, D, layout=layout), abi.field(b, n, 0, D, layout=layout)
    names = ["Periodic"] * (2 * D) if periodic else ["NoSlip"] * (2 * D)
    g = abi.resolve_bc(lid, [abi.BC[x] for x in names])
    F = [1e-6, 0.0, 0.0][:D] if forced else None
    cands = candidates or (CANDIDATES_3D if D == 3 else CANDIDATES_2D)
    out = []
    stream = torch.cuda.current_stream()
    for cfg in cands:
        abi.set_launch(lid, *cfg)
        for _ in range(3):
            abi.step(fa, fb, g, 1.25, F, stream)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        cur, nxt = fa, fb
        for _ in range(steps):
            abi.step(cur, nxt, g, 1.25, F, stream)
            cur, nxt = nxt, cur
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        cells = n[0] * n[1] * n[2]
        out.append({"cfg": list(cfg), "ms": ms, "mlups": cells / ms / 1e3,
                    "GBs": cells * 2 * Q * 8 / ms / 1e6})
        print(json.dumps({"lattice": lattice, "n": n, "layout": layout, **out[-1]}), flush=True)
    out.sort(key=lambda r: r["ms"])
    if apply:
        abi.set_launch(lid, *out[0]["cfg"])
    else:
        abi.set_launch(lid, None)
    return out


CANDIDATES_DEV = [(bx, by, nc, nt) for bx, by in ((256, 1), (128, 2), (64, 4), (64, 2), (128, 1), (64, 1))
                  for nc, nt in ((1, 0), (2, 0), (2, 1))]


def tune_step_dev(lattice: str, dtype: str, n, forced: bool = False, steps: int = 20,
                  candidates=None, apply: bool = True, dev="cuda:0", layout: str = "row"):
    """The same search for reduced storage (FP32 / FP16 / BF16 deviations):
    block shape, one cell or two x-adjacent cells per thread, non-temporal
    loads of the pairs."""
    import torch
    lid, did = abi.LATTICE[lattice], abi.DTYPE[dtype]
    D, Q = abi.LATTICE_DQ[lid]
    n = list(n) + [1] * (3 - len(n))
    tdt = {"f32": torch.float32, "f16dev": torch.float16, "bf16dev": torch.bfloat16}[dtype]
    shape = abi.row_shape(Q, n, 0, D) if layout == "row" else (Q, n[2], n[1], n[0])
    a = torch.zeros(shape, dtype=tdt, device=dev)
    b = torch.zeros_like(a)
    fa = abi.field(a, n, 0, D, did, layout=layout)
    fb = abi.field(b, n, 0, D, did, layout=layout)
    g = abi.resolve_bc(lid, [abi.BC["Periodic"]] * (2 * D))
    F = [1e-6, 0.0, 0.0][:D] if forced else None
    out = []
    stream = torch.cuda.current_stream()
    for cfg in candidates or CANDIDATES_DEV:
        abi.set_launch_dev(lid, did, *cfg)
        for _ in range(3):
            abi.step(fa, fb, g, 1.25, F, stream)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        cur, nxt = fa, fb
        for _ in range(steps):
            abi.step(cur, nxt, g, 1.25, F, stream)
            cur, nxt = nxt, cur
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        cells = n[0] * n[1] * n[2]
        out.append({"cfg": list(cfg), "ms": ms, "mlups": cells / ms / 1e3,
                    "GBs": cells * 2 * Q * a.element_size() / ms / 1e6})
        print(json.dumps({"lattice": lattice, "dtype": dtype, "n": n, "layout": layout, **out[-1]}),
              flush=True)
    out.sort(key=lambda r: r["ms"])
    if apply:
        abi.set_launch_dev(lid, did, *out[0]["cfg"])
    else:
        abi.set_launch_dev(lid, did, None)
    return out


def copy_ceiling(nbytes: int = 2 * 1024 ** 3, reps: int = 10, dev="cuda:0"):
    """GB/s (read + write) of mfa_stream_copy, plain and non-temporal."""
    import torch
    src = torch.empty(nbytes // 8, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    src.fill_(1.0)
    res = {}
    for nt in (0, 1):
        for _ in range(2):
            abi.stream_copy(src.data_ptr(), dst.data_ptr(), nbytes, nt)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            abi.stream_copy(src.data_ptr(), dst.data_ptr(), nbytes, nt)
        e1.record()
        torch.cuda.synchronize()
        res["nt" if nt else "plain"] = 2 * nbytes / (e0.elapsed_time(e1) / reps * 1e-3) / 1e9
    return res


if __name__ == "__main__":
    # python -m muphfasa_amd.autotune LATTICE [DTYPE] NX NY [NZ] [--layout component]
    args = sys.argv[1:]
    layout = "row"
    if "--layout" in args:
        i = args.index("--layout")
        layout = args[i + 1]
        del args[i:i + 2]
    lat = args.pop(0) if args else "D3Q19"
    dtype = args.pop(0) if args and not args[0].isdigit() else "f64"
    size = [int(v) for v in args] or [256, 256, 256]
    print(json.dumps({"copy_ceiling_GBs": copy_ceiling()}), flush=True)
    if dtype == "f64":
        best = tune_step(lat, size, layout=layout)
    else:
        best = tune_step_dev(lat, dtype, size, layout=layout)
    print(json.dumps({"best": best[0]}), flush=True)
