"""Batched ensemble driver: many independent D2Q9 lattices (lid cavities,
optionally with solid obstacles) advanced by one launch per step, graph-
captured in chunks, with per-lattice convergence retirement.

The steady-state criterion is the relative L2 change of the velocity field
over one check interval (default 1 000 steps):
    res = sqrt( sum |u(t) - u(t - K)|^2 / sum |u(t)|^2 )  over fluid cells,
and a lattice retires when res < tol (default 1e-8) or at max_steps.
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

_lib = None


def lib():
    global _lib
    if _lib is None:
        L = abi.lib()          # libmfa_lbm.so carries the ensemble entry points
        vp = ctypes.c_void_p
        ip = ctypes.POINTER(ctypes.c_int)
        dp = ctypes.POINTER(ctypes.c_double)
        lp = ctypes.POINTER(ctypes.c_longlong)
        L.mfa_ensemble_error.restype = ctypes.c_char_p
        L.mfa_ensemble_create.restype = vp
        L.mfa_ensemble_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, ip, dp,
                                          ctypes.POINTER(abi.GeomBC), ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int]
        L.mfa_ensemble_destroy.argtypes = [vp]
        L.mfa_ensemble_info.argtypes = [vp, lp, lp, lp, lp, ip]
        L.mfa_ensemble_set_active.argtypes = [vp, ip]
        L.mfa_ensemble_set_grid.argtypes = [vp, ctypes.c_int]
        L.mfa_ensemble_init.argtypes = [vp, vp, vp, vp, vp, vp]
        L.mfa_ensemble_init_warm.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.mfa_ensemble_init_start.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_int, vp]
        L.mfa_ensemble_step.argtypes = [vp, vp, vp, vp, vp]
        L.mfa_ensemble_step2.argtypes = [vp, vp, vp, vp, vp]
        L.mfa_ensemble_run.argtypes = [vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, vp]
        L.mfa_ensemble_moments.argtypes = [vp, vp, vp, vp, vp, vp]
        L.mfa_ensemble_residual.argtypes = [vp, vp, vp, vp, dp, vp]
        for n in ("mfa_ensemble_info", "mfa_ensemble_set_active", "mfa_ensemble_set_grid",
                  "mfa_ensemble_init", "mfa_ensemble_init_warm", "mfa_ensemble_init_start",
                  "mfa_ensemble_step", "mfa_ensemble_step2",
                  "mfa_ensemble_run",
                  "mfa_ensemble_moments", "mfa_ensemble_residual"):
            getattr(L, n).restype = ctypes.c_int
        _lib = L
    return _lib


START_NEQ, START_AVERAGE = 1, 2   # mfa_ensemble.h MFA_START_*


def _check(rc):
    if rc != 0:
        raise abi.MfaError(rc, lib().mfa_ensemble_error().decode())


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def lid_cavity_bc(u_lid=0.1):
    fu = np.zeros((6, 3))
    fu[3, 0] = u_lid
    return abi.resolve_bc(0, [abi.BC["NoSlip"]] * 3 + [abi.BC["MovingWall"]], [1.0] * 6, fu)


class Ensemble:
    def __init__(self, sizes, omegas, bc=None, dtype="f64", tile=(64, 4), masks=None,
                 dev="cuda:0", grid=0):
        import torch
        self.torch = torch
        self.dev = torch.device(dev)
        self.sizes = [(int(a), int(b)) for a, b in sizes]
        self.nlat = len(self.sizes)
        self.omegas = [float(o) for o in omegas]
        self.dtype = dtype
        self.bc = bc if bc is not None else lid_cavity_bc()
        nx = (ctypes.c_int * self.nlat)(*[s[0] for s in self.sizes])
        ny = (ctypes.c_int * self.nlat)(*[s[1] for s in self.sizes])
        om = (ctypes.c_double * self.nlat)(*self.omegas)
        self.masked = masks is not None
        h = lib().mfa_ensemble_create(0, abi.DTYPE[dtype], self.nlat, nx, ny, om,
                                      ctypes.byref(self.bc), tile[0], tile[1], 1 if self.masked else 0)
        if not h:
            raise abi.MfaError(-1, lib().mfa_ensemble_error().decode())
        self.h = ctypes.c_void_p(h)
        self.tile = tile
        tp = ctypes.c_longlong()
        tc = ctypes.c_longlong()
        po = (ctypes.c_longlong * self.nlat)()
        co = (ctypes.c_longlong * self.nlat)()
        nt = ctypes.c_int()
        _check(lib().mfa_ensemble_info(self.h, ctypes.byref(tp), ctypes.byref(tc), po, co,
                                       ctypes.byref(nt)))
        self.total_pops, self.total_cells = tp.value, tc.value
        self.pop_off, self.cell_off = list(po), list(co)
        self.ntiles = nt.value
        tdt = torch.float64 if dtype == "f64" else torch.float32
        self.fa = torch.empty(self.total_pops, dtype=tdt, device=self.dev)
        self.fb = torch.empty_like(self.fa)
        self.rho = torch.zeros(self.total_cells, dtype=torch.float64, device=self.dev)
        self.vel = torch.zeros(2 * self.total_cells, dtype=torch.float64, device=self.dev)
        self.vel_prev = torch.zeros_like(self.vel)
        self.mask = None
        if self.masked:
            m = np.concatenate([np.asarray(mk, dtype=np.uint8).reshape(-1) for mk in masks])
            assert m.size == self.total_cells
            self.mask = torch.from_numpy(m).to(self.dev)
        self.active = np.ones(self.nlat, dtype=np.int32)
        if grid:
            _check(lib().mfa_ensemble_set_grid(self.h, grid))

    def close(self):
        if getattr(self, "h", None):
            lib().mfa_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stream(self):
        return abi.stream_ptr()

    def init(self, rho=None, vel=None, average=False, neq=False):
        """f = feq(rho, u) (rest state when both are None).  neq=True adds the
        first-order Chapman-Enskog non-equilibrium part of u
        (mfa_ensemble_init_start, MFA_START_NEQ); average=True then takes one
        step and the mean of the two consecutive states (the warm start of
        mfa_ensemble_init_warm)."""
        if neq:
            flags = START_NEQ | (START_AVERAGE if average else 0)
            _check(lib().mfa_ensemble_init_start(self.h, _p(self.fa), _p(self.fb), _p(self.mask),
                                                 _p(rho), _p(vel), flags, self.stream()))
        elif average:
            _check(lib().mfa_ensemble_init_warm(self.h, _p(self.fa), _p(self.fb), _p(self.mask),
                                                _p(rho), _p(vel), self.stream()))
        else:
            _check(lib().mfa_ensemble_init(self.h, _p(self.fa), _p(self.mask), _p(rho), _p(vel),
                                           self.stream()))

    def set_grid(self, blocks):
        _check(lib().mfa_ensemble_set_grid(self.h, blocks))

    def set_active(self, active):
        self.active = np.asarray(active, dtype=np.int32).copy()
        arr = (ctypes.c_int * self.nlat)(*self.active.tolist())
        _check(lib().mfa_ensemble_set_active(self.h, arr))
        nt = ctypes.c_int()
        _check(lib().mfa_ensemble_info(self.h, None, None, None, None, ctypes.byref(nt)))
        self.ntiles = nt.value

    def step(self, pair=False):
        """One launch fa -> fb, then swap: one step, or two with pair=True."""
        fn = lib().mfa_ensemble_step2 if pair else lib().mfa_ensemble_step
        _check(fn(self.h, _p(self.fa), _p(self.fb), _p(self.mask), self.stream()))
        self.fa, self.fb = self.fb, self.fa

    def run(self, nsteps, graph=True):
        """nsteps steps (state stays in fa when nsteps is even)."""
        _check(lib().mfa_ensemble_run(self.h, _p(self.fa), _p(self.fb), _p(self.mask), nsteps,
                                      1 if graph else 0, self.stream()))
        if nsteps % 2:
            self.fa, self.fb = self.fb, self.fa

    def moments(self, out_vel=None):
        _check(lib().mfa_ensemble_moments(self.h, _p(self.fa), _p(self.mask), _p(self.rho),
                                          _p(out_vel if out_vel is not None else self.vel),
                                          self.stream()))

    def residual(self):
        """Relative L2 change of u since the last call (per lattice)."""
        self.moments(self.vel)
        out = (ctypes.c_double * (2 * self.nlat))()
        _check(lib().mfa_ensemble_residual(self.h, _p(self.vel), _p(self.vel_prev),
                                           _p(self.mask), out, self.stream()))
        self.vel_prev.copy_(self.vel)
        a = np.array(out[0::2])
        b = np.array(out[1::2])
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.sqrt(a / b)

    def lattice_fields(self, l):
        nx, ny = self.sizes[l]
        c0 = self.cell_off[l]
        n = nx * ny
        rho = self.rho[c0:c0 + n].cpu().numpy().reshape(ny, nx)
        u = self.vel[c0:c0 + n].cpu().numpy().reshape(ny, nx)
        v = self.vel[self.total_cells + c0:self.total_cells + c0 + n].cpu().numpy().reshape(ny, nx)
        return rho, np.stack([u, v])

    def lattice_populations(self, l):
        """Populations of lattice l as (9, 1, ny, nx), component-major like the
        oracle's fabs (the engine stores them row-interleaved, [y][q][x])."""
        nx, ny = self.sizes[l]
        p0 = self.pop_off[l]
        rows = self.fa[p0:p0 + 9 * nx * ny].double().cpu().numpy().reshape(ny, 9, nx)
        return np.ascontiguousarray(rows.transpose(1, 0, 2)).reshape(9, 1, ny, nx)


def run_to_steady(ens: Ensemble, check=1000, tol=1e-8, max_steps=200000, graph=True, log=None,
                  max_seconds=None):
    """Advance until every lattice's residual < tol or max_steps (or, if
    given, max_seconds of wall time).  Returns a dict with per-lattice
    steps/residual/status and aggregate MLUPS."""
    torch = ens.torch
    steps = np.zeros(ens.nlat, dtype=np.int64)
    res = np.full(ens.nlat, np.inf)
    status = np.array(["running"] * ens.nlat, dtype=object)
    cells = np.array([a * b for a, b in ens.sizes], dtype=np.int64)
    ens.moments(ens.vel_prev)
    t_gpu = 0.0
    lu = 0
    done_steps = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    while ens.active.any() and done_steps < max_steps:
        if max_seconds is not None and time.perf_counter() - t0 > max_seconds:
            break
        k = min(check, max_steps - done_steps)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        ens.run(k, graph=graph)
        e1.record()
        r = ens.residual()
        t_gpu += e0.elapsed_time(e1) * 1e-3
        act = ens.active.astype(bool)
        lu += int(cells[act].sum()) * k
        done_steps += k
        steps[act] = done_steps
        res[act] = r[act]
        newly = act & ((r < tol) | ~np.isfinite(r))
        status[newly & (r < tol)] = "converged"
        status[newly & ~np.isfinite(r)] = "diverged"
        if newly.any():
            a = ens.active.copy()
            a[newly] = 0
            ens.set_active(a)
        if log:
            log(done_steps, int(ens.active.sum()), float(np.nanmax(np.where(act, r, np.nan))))
    status[status == "running"] = "max_steps"
    wall = time.perf_counter() - t0
    return {"steps": steps, "residual": res, "status": status, "gpu_s": t_gpu, "wall_s": wall,
            "lattice_updates": lu, "mlups_gpu": lu / t_gpu / 1e6 if t_gpu else 0.0,
            "mlups_wall": lu / wall / 1e6 if wall else 0.0}


def cavity_ensemble_params(n=512, seed=2024, u_lid=0.1, sizes=range(64, 257, 32),
                           re_range=(50.0, 1000.0)):
    """Configuration 3: side N uniform from {64,96,...,256}, Re log-uniform,
    lid u=0.1, tau = 0.5 + 3*u_lid*N/Re."""
    rng = np.random.default_rng(seed)
    N = rng.choice(list(sizes), size=n)
    Re = np.exp(rng.uniform(math.log(re_range[0]), math.log(re_range[1]), size=n))
    tau = 0.5 + 3.0 * u_lid * N / Re
    return N.astype(int), Re, tau


TILE_CANDIDATES = [(64, 4), (64, 2), (64, 1), (128, 2), (128, 1), (256, 1), (32, 8), (32, 2),
                   (16, 16), (16, 4)]


def tune_tiles(sizes, omegas, dtype="f64", steps=200, candidates=None, masks=None):
    """Time each tile shape (and a persistent grid of 4 blocks per CU versus
    one block per tile) on the given ensemble; returns results sorted by
    time."""
    import torch
    out = []
    for tile in candidates or TILE_CANDIDATES:
        if (tile[0] * tile[1]) % 64:
            continue
        for grid in (0, 1024):
            ens = Ensemble(sizes, omegas, dtype=dtype, tile=tile, masks=masks, grid=grid)
            ens.init()
            ens.run(10, graph=False)
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            ens.run(steps, graph=True)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / steps
            cells = sum(a * b for a, b in sizes)
            out.append({"tile": list(tile), "grid": grid, "ms_per_step": ms,
                        "mlups": cells / ms / 1e3})
            print(json.dumps(out[-1]), flush=True)
            ens.close()
            del ens
    out.sort(key=lambda r: r["ms_per_step"])
    return out


def compare_storage(ens, r, sizes, omegas, a, tile, grid, Re, log):
    """Runs the same cavities with storage a.compare, each for exactly the
    steps its --dtype run took, and prints the distance of its fields from
    the --dtype run's: max over cells of |u - u_ref| / U per cavity (median,
    90th percentile and max over the cavities, by Re band) and the relative
    L2 distance of u."""
    torch = ens.torch
    ens.set_active(np.ones(ens.nlat, np.int32))
    ens.moments()
    ref_vel = ens.vel.clone()
    cmp = Ensemble(sizes, omegas, dtype=a.compare, tile=tile, grid=grid)
    cmp.init()
    done = 0
    steps = r["steps"]
    for target in sorted(set(int(x) for x in steps)):
        if target > done:
            cmp.run(target - done, graph=not a.no_graph)
            done = target
        act = cmp.active.copy()
        act[steps == target] = 0
        # retire the cavities whose reference run stopped here
        if (act != cmp.active).any():
            cmp.set_active(act)
        log(done, int(cmp.active.sum()), float("nan"))
    cmp.set_active(np.ones(cmp.nlat, np.int32))
    cmp.moments()
    torch.cuda.synchronize()
    cells = [x * y for x, y in sizes]
    tc = sum(cells)
    d = (cmp.vel - ref_vel).abs() / 0.1
    dmax = np.array([float(v.max()) for v in torch.maximum(d[:tc], d[tc:]).split(cells)])
    num = [(float((x * x).sum()) + float((y * y).sum())) for x, y in
           zip((cmp.vel - ref_vel)[:tc].split(cells), (cmp.vel - ref_vel)[tc:].split(cells))]
    den = [(float((x * x).sum()) + float((y * y).sum())) for x, y in
           zip(ref_vel[:tc].split(cells), ref_vel[tc:].split(cells))]
    l2 = np.sqrt(np.array(num) / np.maximum(np.array(den), 1e-300))
    ok = r["status"] == "converged"
    bands = []
    for lo, hi in ((50, 100), (100, 200), (200, 400), (400, 1000.01)):
        sel = (Re >= lo) & (Re < hi) & np.isfinite(dmax)
        if sel.any():
            bands.append({"Re": [lo, min(hi, 1000)], "cavities": int(sel.sum()),
                          "max_du_over_U_median": float(np.median(dmax[sel])),
                          "max_du_over_U_max": float(dmax[sel].max())})
    fin = np.isfinite(dmax)
    print(json.dumps({"compare": a.compare, "against": a.dtype, "cavities": int(fin.sum()),
                      "max_du_over_U": {"median": float(np.median(dmax[fin])),
                                        "p90": float(np.percentile(dmax[fin], 90)),
                                        "max": float(dmax[fin].max())},
                      "converged_reference_max_du_over_U_median":
                          float(np.median(dmax[ok & fin])) if (ok & fin).any() else None,
                      "rel_l2_u": {"median": float(np.median(l2[fin])), "max": float(l2[fin].max())},
                      "bands": bands}), flush=True)


def main():
    import argparse
    import torch
    p = argparse.ArgumentParser()
    p.add_argument("--n", type=int, default=512)
    p.add_argument("--dtype", default="f64")
    p.add_argument("--max-steps", type=int, default=200000)
    p.add_argument("--check", type=int, default=1000)
    p.add_argument("--tol", type=float, default=1e-8)
    p.add_argument("--tune", action="store_true")
    p.add_argument("--tile", default="")
    p.add_argument("--grid", type=int, default=0)
    p.add_argument("--no-graph", action="store_true")
    p.add_argument("--out", default="")
    p.add_argument("--compare", default="",
                   help="also run this storage type (e.g. f32) on the same cavities and "
                        "report its steady fields against those of --dtype")
    a = p.parse_args()
    N, Re, tau = cavity_ensemble_params(a.n)
    sizes = [(int(x), int(x)) for x in N]
    omegas = 1.0 / tau
    tile = (64, 4)
    grid = a.grid
    if a.tune:
        best = tune_tiles(sizes, omegas, a.dtype)[0]
        tile, grid = tuple(best["tile"]), best["grid"]
        print(json.dumps({"tuned": best}), flush=True)
    if a.tile:
        tile = tuple(int(v) for v in a.tile.split("x"))
    ens = Ensemble(sizes, omegas, dtype=a.dtype, tile=tile, grid=grid)
    ens.init()
    t0 = time.time()

    def log(s, nact, rmax):
        if s % (10 * a.check) == 0:
            print(json.dumps({"step": s, "active": nact, "max_res": rmax,
                              "t": round(time.time() - t0, 1)}), flush=True)
    r = run_to_steady(ens, a.check, a.tol, a.max_steps, graph=not a.no_graph, log=log)
    if a.compare:
        compare_storage(ens, r, sizes, omegas, a, tile, grid, Re, log)
    summary = {"config": "batched D2Q9 lid cavities", "n_lattices": a.n, "dtype": a.dtype,
               "tile": list(tile), "grid": grid, "tol": a.tol, "check": a.check,
               "max_steps": a.max_steps,
               "converged": int((r["status"] == "converged").sum()),
               "diverged": int((r["status"] == "diverged").sum()),
               "hit_max_steps": int((r["status"] == "max_steps").sum()),
               "lattice_updates": r["lattice_updates"], "gpu_s": round(r["gpu_s"], 3),
               "wall_s": round(r["wall_s"], 3), "aggregate_mlups_gpu": round(r["mlups_gpu"], 1),
               "aggregate_mlups_wall": round(r["mlups_wall"], 1),
               "bytes_per_lu": 2 * 9 * (8 if a.dtype == "f64" else 4)}
    print(json.dumps(summary), flush=True)
    if a.out:
        np.savez_compressed(a.out, N=N, Re=Re, tau=tau, steps=r["steps"], residual=r["residual"],
                            status=r["status"].astype(str))


if __name__ == "__main__":
    main()
