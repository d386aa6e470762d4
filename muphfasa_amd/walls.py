"""3-D boxes with walls: the two-step pass (mfa_step2) against two single
steps (mfa_step), FP64, row-interleaved populations.

Boundary sets (lattice units):
  lid         NoSlip on every face but y-hi, a MovingWall at u = (0.1, 0, 0.03)
  noslip      NoSlip on every face
  xper_walls  periodic in x, NoSlip at y-lo, z-lo and z-hi, MovingWall at y-hi
              (u = (0.05, 0, 0))

``python -m muphfasa_amd.walls --n 256 --lattice D3Q19`` prints, per set, the
ms per step of both paths and whether the pair equals the two single steps
bit for bit.
"""
from __future__ import annotations

import argparse
import json

import numpy as np

from . import abi

SETS = {
    "lid": (["NoSlip", "NoSlip", "NoSlip", "MovingWall", "NoSlip", "NoSlip"], {3: [0.1, 0.0, 0.03]}),
    "noslip": (["NoSlip"] * 6, {}),
    "xper_walls": (["Periodic", "Periodic", "NoSlip", "MovingWall", "NoSlip", "NoSlip"], {3: [0.05, 0.0, 0.0]}),
}


def geometry(lattice: str, name: str):
    faces, vel = SETS[name]
    fu = np.zeros((6, 3))
    for i, u in vel.items():
        fu[i] = u
    return abi.resolve_bc(abi.LATTICE[lattice], [abi.BC[f] for f in faces], [1.0] * 6, fu)


def run(n=256, lattice="D3Q19", sets=("lid", "noslip", "xper_walls"), reps=30, tau=0.8, dev="cuda:0"):
    import torch
    Q = {"D3Q15": 15, "D3Q19": 19, "D3Q27": 27}[lattice]
    LI = abi.LATTICE[lattice]
    nn = [n, n, n]
    gen = torch.Generator(device=dev)
    gen.manual_seed(7)
    rho = torch.ones((1, n, n, n), dtype=torch.float64, device=dev)
    vel = (torch.rand((3, n, n, n), dtype=torch.float64, device=dev, generator=gen) * 2 - 1) * 1e-3
    shape = abi.row_shape(Q, nn, 0, 3)
    ta = torch.empty(shape, dtype=torch.float64, device=dev)
    tb = torch.empty_like(ta)
    tc = torch.empty_like(ta)
    fa, fb, fc = [abi.field(t, nn, 0, 3, layout="row") for t in (ta, tb, tc)]
    abi.init_equilibrium(abi.field(rho, nn, 0, 3), abi.field(vel, nn, 0, 3), fa, LI)
    del rho, vel
    omega = 1.0 / tau
    out = {"lattice": lattice, "n": n}

    def timed(fn):
        for _ in range(5):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps / 2          # ms per step
    for name in sets:
        g = geometry(lattice, name)
        abi.step(fa, fc, g, omega)
        abi.step(fc, fb, g, omega)
        torch.cuda.synchronize()
        ref = tb.clone()
        abi.step2(fa, fb, g, omega)
        torch.cuda.synchronize()
        same = bool(torch.equal(tb.view(torch.int64), ref.view(torch.int64)))
        out[name] = {"two_single_steps_ms": round(timed(lambda: (abi.step(fa, fc, g, omega),
                                                                  abi.step(fc, fb, g, omega))), 4),
                     "pair_ms": round(timed(lambda: abi.step2(fa, fb, g, omega)), 4),
                     "bit_identical": same}
        print(json.dumps({name: out[name]}), flush=True)
    return out


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--n", type=int, default=256)
    p.add_argument("--lattice", default="D3Q19", choices=["D3Q15", "D3Q19", "D3Q27"])
    p.add_argument("--reps", type=int, default=30)
    p.add_argument("--sets", default="lid,noslip,xper_walls",
                   help="comma-separated boundary sets (" + ", ".join(SETS) + ")")
    a = p.parse_args()
    print(json.dumps(run(a.n, a.lattice, sets=a.sets.split(","), reps=a.reps)), flush=True)


if __name__ == "__main__":
    main()
