"""A periodic D3Q19 domain chopped into boxes of at most max_grid_size cells
per side, stepped through the multi-box level (include/mfa_level.h), against
the same domain as one box.

The chopping matches the role of AMReX's ``amr.max_grid_size`` for the
reference solver (default 64, so a 256^3 domain runs as 64 boxes of 64^3):
each dimension is cut into ceil(N / max_grid_size) pieces whose sizes differ
by at most one cell.  A step of the level is one ghost exchange
(``mfa_level_fill_boundary``), the physical boundary fill (nothing on a
periodic box) and the step of every box through its ghosts in one launch
(``mfa_level_step``), or the direct step that reads neighbouring boxes' cells
in place (``mfa_level_step_direct``, periodic grids of boxes); the sequence of
two steps is captured once in a HIP graph and replayed.

``python -m muphfasa_amd.level --n 256 --max-grid 64 --steps 200`` prints the
MLUPS of the level and of the single box, and whether their final states are
bit-identical.
"""
from __future__ import annotations

import argparse
import json
import math

from . import abi


def chop(n: int, max_grid: int):
    """Cut [0, n) into ceil(n / max_grid) nearly equal pieces: (lo, size)."""
    k = max(1, math.ceil(n / max_grid))
    base, extra = divmod(n, k)
    out, lo = [], 0
    for i in range(k):
        m = base + (1 if i < extra else 0)
        out.append((lo, m))
        lo += m
    return out


def boxes(domain, max_grid):
    lo, n = [], []
    for z0, nz in chop(domain[2], max_grid):
        for y0, ny in chop(domain[1], max_grid):
            for x0, nx in chop(domain[0], max_grid):
                lo.append([x0, y0, z0])
                n.append([nx, ny, nz])
    return lo, n


class BoxLevel:
    """Two population buffers of every box (ngrow 1), bound to level slots
    0 and 1.  layout "component": one plane per population (the MultiFab's
    order); "row": the Q rows of one (y, z) side by side."""

    def __init__(self, lattice: str, domain, max_grid: int, g, dev="cuda:0", layout="component", ngrow=1):
        import torch
        self.L = abi.LATTICE[lattice]
        self.Q = {"D3Q15": 15, "D3Q19": 19, "D3Q27": 27}[lattice]
        self.domain = list(domain)
        self.lo, self.n = boxes(self.domain, max_grid)
        self.g = g
        per = [int(g.periodic[d]) for d in range(3)]
        self.level = abi.Level(self.L, self.domain, per, self.lo, self.n)
        self.g0 = ngrow
        e = 2 * ngrow
        if layout == "component":
            shape = lambda m: (self.Q, m[2] + e, m[1] + e, m[0] + e)
            self.t = [[torch.full(shape(m), float("nan"), dtype=torch.float64, device=dev) for m in self.n]
                      for _ in range(2)]
            self.cv = self.t
        else:
            shape = lambda m: (m[2] + e, m[1] + e, self.Q, m[0] + e)
            self.t = [[torch.full(shape(m), float("nan"), dtype=torch.float64, device=dev) for m in self.n]
                      for _ in range(2)]
            self.cv = [[t.permute(2, 0, 1, 3) for t in ts] for ts in self.t]   # (Q, z, y, x) views
        self.f = [[abi.field(t, self.n[b], ngrow, 3, layout=layout) for b, t in enumerate(ts)] for ts in self.t]
        self.level.bind(0, self.f[0])
        self.level.bind(1, self.f[1])

    def _valid(self, slot, b):
        g, m = self.g0, self.n[b]
        return self.cv[slot][b][:, g:g + m[2], g:g + m[1], g:g + m[0]]

    def scatter(self, full, slot=0):
        """Copy a (Q, nz, ny, nx) tensor of the whole domain into the boxes."""
        for b, (l, m) in enumerate(zip(self.lo, self.n)):
            self._valid(slot, b)[...] = full[:, l[2]:l[2] + m[2], l[1]:l[1] + m[1], l[0]:l[0] + m[0]]

    def gather(self, full, slot=0):
        for b, (l, m) in enumerate(zip(self.lo, self.n)):
            full[:, l[2]:l[2] + m[2], l[1]:l[1] + m[1], l[0]:l[0] + m[0]] = self._valid(slot, b)

    def step(self, src: int, omega: float, force=None, direct=False):
        """One step from slot src into the other slot: the exchange, the
        physical boundary fill and the step of every box, or (direct=True)
        the step that reads neighbouring boxes' cells in place."""
        if direct:
            self.level.step_direct(src, 1 - src, omega, force)
            return
        self.level.fill_boundary(src)
        self.level.apply_bc(src, self.g)
        self.level.step(src, 1 - src, omega, force)

    def close(self):
        self.level.close()


def run(n=256, max_grid=64, steps=200, tau=0.8, force=(1e-6, 0.0, 0.0), dev="cuda:0", layout="component"):
    import torch
    lat = "D3Q19"
    Q = 19
    dom = [n, n, n]
    g = abi.resolve_bc(abi.LATTICE[lat], [abi.BC["Periodic"]] * 6)
    omega = 1.0 / tau
    F = list(force)
    # initial state: feq of rho = 1 and seeded small velocities
    torch.manual_seed(2024)
    rho = torch.ones((1, n, n, n), dtype=torch.float64, device=dev)
    vel = (torch.rand((3, n, n, n), dtype=torch.float64, device=dev) - 0.5) * 2e-3
    full = torch.empty((Q, n, n, n), dtype=torch.float64, device=dev)
    abi.init_equilibrium(abi.field(rho, dom, 0, 3), abi.field(vel, dom, 0, 3), abi.field(full, dom, 0, 3),
                         abi.LATTICE[lat], F)
    del rho, vel
    steps -= steps % 2
    out = {"workload": f"D3Q19 periodic {n}^3 FP64 forced, boxes of at most {max_grid}^3",
           "steps": steps, "layout": layout}
    # one box
    if layout == "component":
        a, b = full.clone(), torch.empty_like(full)
        av = a
    else:
        a = full.permute(1, 2, 0, 3).contiguous()          # (z, y, Q, x)
        b = torch.empty_like(a)
        av = a.permute(2, 0, 1, 3)
    fa, fb = abi.field(a, dom, 0, 3, layout=layout), abi.field(b, dom, 0, 3, layout=layout)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        abi.step(fa, fb, g, omega, F)
        abi.step(fb, fa, g, omega, F)
    torch.cuda.synchronize()
    av.copy_(full)
    gr = torch.cuda.CUDAGraph()
    with torch.cuda.graph(gr):
        abi.step(fa, fb, g, omega, F)
        abi.step(fb, fa, g, omega, F)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps // 2):
        gr.replay()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    out["one_box"] = {"boxes": 1, "ms_per_step": round(ms / steps, 4),
                      "mlups": round(n ** 3 * steps / (ms * 1e-3) / 1e6, 1)}
    single = av.contiguous()
    del a, b, av, gr
    # the level, with the exchange and with the direct step
    lv = BoxLevel(lat, dom, max_grid, g, dev, layout)
    got = torch.empty_like(full)
    for direct in ((False, True) if lv.level.direct_available else (False,)):
        lv.scatter(full, 0)
        with torch.cuda.stream(s):
            lv.step(0, omega, F, direct)
            lv.step(1, omega, F, direct)
        torch.cuda.synchronize()
        lv.scatter(full, 0)
        gl = torch.cuda.CUDAGraph()
        with torch.cuda.graph(gl):
            lv.step(0, omega, F, direct)
            lv.step(1, omega, F, direct)
        e0.record()
        for _ in range(steps // 2):
            gl.replay()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        key = "level_direct" if direct else "level"
        out[key] = {"boxes": len(lv.n), "ms_per_step": round(ms / steps, 4),
                    "mlups": round(n ** 3 * steps / (ms * 1e-3) / 1e6, 1)}
        if not direct:
            out[key]["ghost_cells_per_exchange"] = lv.level.ghost_cells
        lv.gather(got, 0)
        out[key]["bit_identical"] = bool(torch.equal(got.view(torch.int64), single.view(torch.int64)))
        del gl
    direct_ok = lv.level.direct_available
    lv.close()
    if direct_ok:
        # the direct step on boxes stored without a ghost layer
        lv = BoxLevel(lat, dom, max_grid, g, dev, layout, ngrow=0)
        lv.scatter(full, 0)
        with torch.cuda.stream(s):
            lv.step(0, omega, F, True)
            lv.step(1, omega, F, True)
        torch.cuda.synchronize()
        lv.scatter(full, 0)
        gl = torch.cuda.CUDAGraph()
        with torch.cuda.graph(gl):
            lv.step(0, omega, F, True)
            lv.step(1, omega, F, True)
        e0.record()
        for _ in range(steps // 2):
            gl.replay()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        lv.gather(got, 0)
        out["level_direct_no_ghosts"] = {
            "boxes": len(lv.n), "ms_per_step": round(ms / steps, 4),
            "mlups": round(n ** 3 * steps / (ms * 1e-3) / 1e6, 1),
            "bit_identical": bool(torch.equal(got.view(torch.int64), single.view(torch.int64)))}
        del gl
        lv.close()
    out["bit_identical"] = all(out[k]["bit_identical"] for k in ("level", "level_direct", "level_direct_no_ghosts")
                               if k in out)
    return out


def tune_direct(n=256, max_grid=64, steps=100, tau=0.8, force=(1e-6, 0.0, 0.0), dev="cuda:0",
                layout="component"):
    """Times the direct step for every launch shape (cells per thread along z
    1, 2, 4; plain, non-temporal, or non-temporal c_x = 0 loads) on the
    chopped periodic D3Q19 domain; every shape's state must equal the
    default's bit for bit.  Returns the rows sorted by time."""
    import torch
    dom = [n, n, n]
    g = abi.resolve_bc(abi.LATTICE["D3Q19"], [abi.BC["Periodic"]] * 6)
    omega = 1.0 / tau
    F = list(force)
    torch.manual_seed(2024)
    rho = torch.ones((1, n, n, n), dtype=torch.float64, device=dev)
    vel = (torch.rand((3, n, n, n), dtype=torch.float64, device=dev) - 0.5) * 2e-3
    full = torch.empty((19, n, n, n), dtype=torch.float64, device=dev)
    abi.init_equilibrium(abi.field(rho, dom, 0, 3), abi.field(vel, dom, 0, 3), abi.field(full, dom, 0, 3),
                         abi.LATTICE["D3Q19"], F)
    del rho, vel
    lv = BoxLevel("D3Q19", dom, max_grid, g, dev, layout)
    rows, ref = [], None
    got = torch.empty_like(full)
    for nc in (1, 2, 4):
        for nt in (0, 1, 2):
            lv.level.set_direct_launch(nc, nt)
            lv.scatter(full, 0)
            for _ in range(2):
                lv.step(0, omega, F, True)
                lv.step(1, omega, F, True)
            torch.cuda.synchronize()
            lv.gather(got, 0)
            same = True if ref is None else bool(torch.equal(got.view(torch.int64), ref.view(torch.int64)))
            if ref is None:
                ref = got.clone()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(steps // 2):
                lv.step(0, omega, F, True)
                lv.step(1, omega, F, True)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / (steps - steps % 2)
            rows.append({"cells_per_thread": nc, "nontemporal": nt, "ms_per_step": round(ms, 4), "same": same})
            print(json.dumps(rows[-1]), flush=True)
    lv.level.set_direct_launch(2, 0)
    lv.close()
    return sorted(rows, key=lambda r: r["ms_per_step"])


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--n", type=int, default=256)
    p.add_argument("--max-grid", type=int, default=64)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--layout", default="component", choices=["component", "row"])
    p.add_argument("--tune-direct", action="store_true",
                   help="time every launch shape of the direct step instead")
    a = p.parse_args()
    if a.tune_direct:
        print(json.dumps({"best": tune_direct(a.n, a.max_grid, a.steps, layout=a.layout)[0]}), flush=True)
        return
    print(json.dumps(run(a.n, a.max_grid, a.steps, layout=a.layout)), flush=True)


if __name__ == "__main__":
    main()
