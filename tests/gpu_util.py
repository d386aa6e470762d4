"""Shared helpers for the GPU parity tests: move oracle-layout fabs to and from
device memory and build mfa_field views over them."""
from __future__ import annotations

import numpy as np

from muphfasa_amd import abi
from oracle import oracle as O


def to_dev(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def to_host(t):
    return t.detach().cpu().numpy()


def ghosted(L, n, dev, host_fab=None):
    """Device copy of an oracle fab (ngrow 1) and its mfa_field."""
    fab = host_fab if host_fab is not None else O.new_fab(L, n)
    t = to_dev(fab, dev)
    return t, abi.field(t, n, 1, L.D)


def valid_only(L, n, dev, host_fab=None):
    """Device array of the valid cells only (ngrow 0)."""
    import torch
    nn = O.n3(n)
    if host_fab is None:
        t = torch.full((L.Q, nn[2], nn[1], nn[0]), float("nan"), dtype=torch.float64, device=dev)
    else:
        t = to_dev(np.ascontiguousarray(O.valid(L, host_fab)), dev)
    return t, abi.field(t, n, 0, L.D)


def scalar_fields(L, n, dev, rho=None, vel=None):
    import torch
    nn = O.n3(n)
    r = torch.zeros((1, nn[2], nn[1], nn[0]), dtype=torch.float64, device=dev) if rho is None \
        else to_dev(np.asarray(rho).reshape(1, nn[2], nn[1], nn[0]), dev)
    v = torch.zeros((L.D, nn[2], nn[1], nn[0]), dtype=torch.float64, device=dev) if vel is None \
        else to_dev(np.asarray(vel).reshape(L.D, nn[2], nn[1], nn[0]), dev)
    return r, abi.field(r, n, 0, L.D), v, abi.field(v, n, 0, L.D)


def random_state(L, n, seed, amp=0.05):
    """Valid cells near equilibrium with seeded noise; ghosts NaN."""
    rng = np.random.default_rng(seed)
    f = O.new_fab(L, n)
    v = O.valid(L, f)
    v[...] = L.w[:, None, None, None] * (1.0 + amp * rng.standard_normal(v.shape))
    return f


BC_SETS = {
    "periodic": lambda D: (["Periodic"] * (2 * D), None),
    "noslip": lambda D: (["NoSlip"] * (2 * D), None),
    "lid": lambda D: (["NoSlip", "NoSlip", "NoSlip", "MovingWall"] + ["NoSlip"] * (2 * D - 4),
                      {3: [0.1, 0.0, 0.03 if D == 3 else 0.0]}),
    "xper_walls": lambda D: (["Periodic", "Periodic", "NoSlip", "MovingWall"] +
                             ["NoSlip"] * (2 * D - 4), {3: [0.05, 0.0, 0.0]}),
    "inflow_pressure": lambda D: (["Inflow", "PressureOutlet"] + ["NoSlip"] * (2 * D - 2),
                                  {0: [0.04, 0.0, 0.0]}),
    "inflow_outflow": lambda D: (["Inflow", "Outflow"] + ["Periodic"] * (2 * D - 2),
                                 {0: [0.03, 0.0, 0.0]}),
}


def bc_tables(L, name, rho_out=1.02):
    names, vel = BC_SETS[name](L.D)
    ft = [abi.BC[x] for x in names]
    fr = [1.0] * 6
    for i, x in enumerate(names):
        if x == "PressureOutlet":
            fr[i] = rho_out
    fu = np.zeros((6, 3))
    for i, u in (vel or {}).items():
        fu[i, : L.D] = u[: L.D]
    per = [int(names[2 * d] == "Periodic") for d in range(L.D)] + [0] * (3 - L.D)
    g = abi.resolve_bc(abi.LATTICE[L.name], ft, fr, fu)
    ref = O.resolve_bc(L.D, ft, fr, fu)
    return g, ref, per
