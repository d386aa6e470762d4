"""The reference's analytic regression anchors as plain functions: the closed
forms, the error measures and the observed-order fit its regression scripts
use, plus two runners that produce snapshots of (rho, u) for an inputs text --
one through the C++ host on the GPU, one through the oracle on the CPU.

Closed forms and measures follow
  tests/regression/couette/main.cpp:50-68        (Couette start-up series)
  tests/regression/couette/helper.py:223-226,289-350 (profile error, mode-1 fit)
  tests/regression/couette/convergence.py:45-116 (grid convergence at nu t/H^2)
  tests/regression/tgv2d/TgvAnalytic.H:34-50      (2-D Taylor-Green vortex)
  tests/regression/tgv2d/convergence.py:138-222   (TGV grid convergence)
  tests/regression/tgv2d/plot_tgv2d.py:194-209    (max|u| decay-rate fit)
  tests/regression/poiseuille/main.cpp:85-121,130-229 (window statistics)
  tests/regression/poiseuille/convergence.py:37-60 (Poiseuille grid convergence)
All velocities here are in lattice units unless a name says _phys; the
reference's convergence scripts diff physical plotfiles, so their errors are
the lattice errors times dx/dt.
"""
from __future__ import annotations

import math

import numpy as np

CS2 = 1.0 / 3.0


# ---------------------------------------------------------------------------
# closed forms

def couette(y, H, U0, nu, t, n_terms=200):
    """u(y, t) of the impulsively started plane Couette flow (lower wall at
    rest, upper wall at U0, fluid at rest at t = 0), couette/main.cpp:50-68."""
    y = np.asarray(y, dtype=np.float64)
    u = U0 * y / H
    for n in range(1, n_terms + 1):
        arg = -(n * n) * math.pi ** 2 * nu * t / (H * H)
        if arg < -50:
            break
        sign = 1.0 if n % 2 == 0 else -1.0
        u = u + (2.0 * U0 * sign / (n * math.pi)) * np.sin(n * math.pi * y / H) * math.exp(arg)
    return u


def tgv2d(n, u0, nu, t):
    """(rho, u, v) of the 2-D Taylor-Green vortex on an n x n periodic box,
    one vortex pair per side, cell centres at i + 0.5 (TgvAnalytic.H:34-50).
    Arrays are indexed [y, x]."""
    k = 2.0 * math.pi / n
    c = np.arange(n) + 0.5
    x, y = np.meshgrid(c, c)
    d = math.exp(-2.0 * nu * k * k * t)
    d2 = math.exp(-4.0 * nu * k * k * t)
    u = -u0 * np.cos(k * x) * np.sin(k * y) * d
    v = u0 * np.sin(k * x) * np.cos(k * y) * d
    rho = 1.0 - (u0 * u0 / (4.0 * CS2)) * (np.cos(2 * k * x) + np.cos(2 * k * y)) * d2
    return rho, u, v


def poiseuille(H, F, nu):
    """u(y) = F/(2 nu) y (H - y) at the cell centres y = j + 0.5."""
    y = np.arange(int(H)) + 0.5
    return F / (2.0 * nu) * y * (H - y)


# ---------------------------------------------------------------------------
# measures

def wall_profile(vel, D):
    """The x-velocity averaged over the periodic directions, as a function of
    the wall-normal index (y in 2-D, z in 3-D): what AMReX's faverage returns
    along the last dimension.  vel is (D, nz, ny, nx)."""
    ux = vel[0]
    return ux[0].mean(axis=1) if D == 2 else ux.mean(axis=(1, 2))


def wall_field(vel, D):
    """u_x with the wall-normal index first: (H, lateral cells...)."""
    return vel[0, 0] if D == 2 else vel[0]


def first_mode_amplitude(y_norm, u, U0):
    """helper.py:223-226: projection of u - U0 y on sin(pi y)."""
    dev = u - U0 * y_norm
    return float(np.sum(dev * np.sin(np.pi * y_norm)) * 2.0 / len(y_norm))


def mode1_decay_rate(times, amps):
    """helper.py:338-350: least-squares slope of log|a1|(t), skipping the first
    min(3, valid/4) valid points (higher modes still present there)."""
    amps = np.asarray(amps, dtype=np.float64)
    la = np.array([math.log(abs(a)) if abs(a) > 1e-12 else np.nan for a in amps])
    valid = ~np.isnan(la)
    skip = min(3, int(valid.sum()) // 4)
    idx = np.where(valid)[0][skip:]
    if len(idx) < 2:
        raise AssertionError("too few snapshots to fit a decay rate")
    return float(np.polyfit(np.asarray(times, dtype=np.float64)[idx], la[idx], 1)[0])


def observed_orders(errs, ratio=2.0):
    """convergence.py:105-116: log(e_coarse / e_fine) / log(refinement)."""
    return [math.log(errs[i - 1] / errs[i]) / math.log(ratio) for i in range(1, len(errs))]


def poiseuille_window(rho, vel, D, i_lo, i_hi, H, F, nu):
    """poiseuille/main.cpp:85-121 over the cells i_lo <= i < i_hi: max |u_x -
    u_a|, mean u_x, max transverse |u|, rho spread."""
    ua = poiseuille(H, F, nu)
    if D == 2:
        ux = vel[0, 0, :, i_lo:i_hi]
        shape = (-1, 1)
        tr = np.abs(vel[1, 0, :, i_lo:i_hi]).max()
        r = rho[0, :, i_lo:i_hi]
    else:
        ux = vel[0, :, :, i_lo:i_hi]
        shape = (-1, 1, 1)
        tr = max(np.abs(vel[1, :, :, i_lo:i_hi]).max(), np.abs(vel[2, :, :, i_lo:i_hi]).max())
        r = rho[:, :, i_lo:i_hi]
    err = float(np.abs(ux - ua.reshape(shape)).max())
    return err, float(ux.mean()), float(tr), float(r.max() - r.min())


# ---------------------------------------------------------------------------
# inputs

def parse_inputs(text, overrides=()):
    """key -> list of value strings, the last assignment winning, then the
    overrides (key=value ...)."""
    d = {}
    for line in text.splitlines():
        line = line.split("#", 1)[0].strip()
        if "=" in line:
            k, v = line.split("=", 1)
            d[k.strip()] = v.split()
    for ov in overrides:
        k, v = ov.split("=", 1)
        d[k.strip()] = v.split()
    return d


def step_of(text, overrides=()):
    return int(parse_inputs(text, overrides)["sim.max_step"][0])


# ---------------------------------------------------------------------------
# runners: yield (step, time, rho, vel) after each listed step, the way the
# reference's drivers call advance(); writeOutput(step) and then read
# sim->rho(), sim->vel() and sim->time().

def host_snapshots(text, overrides, steps):
    from muphfasa_amd import host
    sim = host.Simulation(text, list(overrides))
    sim.initialize()
    done = 0
    for s in steps:
        sim.advance(s + 1 - done)
        done = s + 1
        sim.update_macroscopic(s)
        rho, vel = sim.macro()
        yield s, sim.time, rho, vel
    sim.close()


_LATTICE_D = {"D2Q9": 2, "D3Q15": 3, "D3Q19": 3, "D3Q27": 3}


def oracle_snapshots(text, overrides, steps):
    """The same snapshots from the oracle, for unforced runs with Periodic,
    NoSlip and MovingWall faces and the rest or tgv2d start."""
    from muphfasa_amd import host
    from oracle import oracle as O
    p = parse_inputs(text, overrides)
    u = host.inputs_units(text, list(overrides))
    name = p["lbm.lattice_type"][0]
    D = _LATTICE_D[name]
    L = O.Lattice(name)
    n = [int(v) for v in p["grid.ncells"]][:D]
    if any(abs(u[k]) > 0 for k in ("F0", "F1", "F2")):
        raise ValueError("oracle_snapshots runs unforced inputs only")
    lo, hi = p["bc.lo_type"], p["bc.hi_type"]
    faces = []
    for d in range(D):
        faces += [O.BC[lo[d]], O.BC[hi[d]]]
    for t in lo + hi:
        if t not in ("Periodic", "NoSlip", "MovingWall"):
            raise ValueError(f"oracle_snapshots: face type {t} not covered")
    fu = np.zeros((6, 3))
    # buildBCManager (src/Simulation/SimulationSetup.cpp:263-273): only a
    # MovingWall face takes its bc.*_Uwall; NoSlip faces are at rest
    for f, key in enumerate(["bc.lox_Uwall", "bc.hix_Uwall", "bc.loy_Uwall", "bc.hiy_Uwall",
                             "bc.loz_Uwall", "bc.hiz_Uwall"][:2 * D]):
        if key in p and (lo, hi)[f % 2][f // 2] == "MovingWall":
            for c, v in enumerate(p[key][:D]):
                fu[f, c] = float(v) * u["dt"] / u["dx"]
    res = O.resolve_bc(D, faces, np.ones(6), fu)
    periodic = [1 if lo[d] == "Periodic" else 0 for d in range(D)]
    nn = O.n3(n)
    init = p.get("lbm.init", ["rest"])[0]
    if init == "rest":
        rho0 = np.ones((nn[2], nn[1], nn[0]))
        vel0 = np.zeros((D, nn[2], nn[1], nn[0]))
    elif init == "tgv2d":
        ref_u = float(p["units.ref_velocity"][0]) * u["dt"] / u["dx"]
        r, a, b = tgv2d(n[0], ref_u, 0.0, 0.0)
        rho0, vel0 = r[None], np.stack([a, b])[:, None]
    else:
        raise ValueError(f"oracle_snapshots: lbm.init = {init} not covered")
    f, _, _ = O.init_equilibrium(L, rho0, vel0, n)
    O.fill_periodic(L, f, n, periodic + [0] * (3 - D))
    O.apply_bc(L, f, n, periodic + [0] * (3 - D), res)
    done = 0
    for s in steps:
        f = O.run(L, f, n, u["omega"], None, periodic, res, s + 1 - done)
        done = s + 1
        rho, vel = O.moments(L, f, n)
        yield s, float(done), rho, vel


# ---------------------------------------------------------------------------
# the reference's checks, as functions of a runner

def couette_startup(run, text, overrides=(), every=50):
    """plot_couette.py / plot_couette3d.py: returns (max profile error / U0
    over all snapshots but the first, relative error of the fitted mode-1
    decay rate against -pi^2 nu / H^2)."""
    from muphfasa_amd import host
    p = parse_inputs(text, overrides)
    u = host.inputs_units(text, list(overrides))
    D = _LATTICE_D[p["lbm.lattice_type"][0]]
    H = float(p["grid.ncells"][D - 1])
    nu = (u["tau"] - 0.5) / 3.0
    key = "bc.hiy_Uwall" if D == 2 else "bc.hiz_Uwall"
    U0 = float(p[key][0]) * u["dt"] / u["dx"]
    y = np.arange(int(H)) + 0.5
    times, errs, amps = [], [], []
    for s, t, rho, vel in run(text, overrides, range(0, step_of(text, overrides) + 1, every)):
        prof = wall_profile(vel, D)
        ua = couette(y, H, U0, nu, t)
        errs.append(float(np.abs(prof - ua).max()) / U0)
        times.append(s + 1.0)
        amps.append(first_mode_amplitude(y / H, prof, U0))
    rate = mode1_decay_rate(times, amps)
    analytic = -(math.pi ** 2) * nu / H ** 2
    return max(errs[1:] or errs), abs(rate - analytic) / abs(analytic)


def couette_convergence(run, text, walls, tau=0.8, target=0.05):
    """convergence.py / convergence3d.py: the inf-norm u_x error in physical
    units at nu t/H^2 = target, for each wall-normal size in walls, with tau
    pinned and the lateral cells scaled with the base ratio."""
    p = parse_inputs(text)
    nc = [int(v) for v in p["grid.ncells"]]
    D = len(nc)
    nu = (tau - 0.5) * CS2
    errs = []
    from muphfasa_amd import host
    for h in walls:
        lat = [max(1, round(h * c / nc[-1])) for c in nc[:-1]]
        step = max(1, round(target * h * h / nu))
        ov = [f"grid.ncells={' '.join(str(v) for v in lat + [h])}", f"grid.max_grid_size={h}",
              f"lbm.tau={tau}", f"sim.max_step={step}"]
        u = host.inputs_units(text, ov)
        key = "bc.hiy_Uwall" if D == 2 else "bc.hiz_Uwall"
        U0 = float(p[key][0]) * u["dt"] / u["dx"]
        y = np.arange(h) + 0.5
        (_, t, rho, vel), = list(run(text, ov, [step]))
        ua = couette(y, float(h), U0, nu, t)
        shape = (-1,) + (1,) * (D - 1)
        e = float(np.abs(wall_field(vel, D) - ua.reshape(shape)).max()) * u["dx"] / u["dt"]
        errs.append(e)
    return errs


def tgv2d_convergence(run, text, sizes, tau=0.8, u0=0.02, target=0.1):
    """tgv2d/convergence.py:138-222: max over u, v of the inf-norm error in
    physical units at nu k^2 t = target, tau pinned."""
    from muphfasa_amd import host
    nu = (tau - 0.5) * CS2
    errs = []
    for n in sizes:
        k = 2.0 * math.pi / n
        step = max(1, round(target / (nu * k * k)))
        ov = [f"grid.ncells={n} {n}", f"grid.max_grid_size={n}", f"lbm.tau={tau}",
              f"units.ref_velocity={u0}", f"sim.max_step={step}"]
        u = host.inputs_units(text, ov)
        (_, t, rho, vel), = list(run(text, ov, [step]))
        _, ua, va = tgv2d(n, u["u_lat"], nu, t)
        e = max(np.abs(vel[0, 0] - ua).max(), np.abs(vel[1, 0] - va).max())
        errs.append(float(e) * u["dx"] / u["dt"])
    return errs


def tgv2d_decay(run, text, every=200):
    """main.cpp:100-115 velocity error at every snapshot (relative to u0) and
    plot_tgv2d.py:194-209 relative error of the fitted max|u| decay rate
    against -2 nu k^2."""
    from muphfasa_amd import host
    u = host.inputs_units(text)
    p = parse_inputs(text)
    n = int(p["grid.ncells"][0])
    nu = (u["tau"] - 0.5) / 3.0
    k = 2.0 * math.pi / n
    u0 = u["u_lat"]
    errs, times, speed = [], [], []
    for s, t, rho, vel in run(text, (), range(0, step_of(text) + 1, every)):
        _, ua, va = tgv2d(n, u0, nu, t)
        errs.append(max(np.abs(vel[0, 0] - ua).max(), np.abs(vel[1, 0] - va).max()) / u0)
        times.append(t)
        speed.append(np.sqrt(vel[0, 0] ** 2 + vel[1, 0] ** 2).max())
    rate = float(np.polyfit(np.array(times), np.log(np.array(speed)), 1)[0])
    analytic = -2.0 * nu * k * k
    return max(errs), abs(rate - analytic) / abs(analytic)
