"""The reference's analytic regression anchors through the C++ host on the
GPU, with the reference's CI criteria:
  plane Couette start-up, 2-D D2Q9 and 3-D D3Q15/19/27: profile error <= 1 %
    of U0 and mode-1 decay rate within 1 %
    (.github/workflows/couette-regression.yml:77,146);
  Couette grid convergence over 16/32/64 wall-normal cells at tau 0.8:
    observed order >= 1.8 (couette-regression.yml:73,142);
  2-D Taylor-Green vortex: grid convergence over 32/64/128 at order >= 1.8
    (.github/workflows/tgv2d-regression.yml:81) and the decay of the default
    64^2 run against TgvAnalytic.H;
  plane Poiseuille in 3-D, periodic (poiseuille/inputs3d) and inflow /
    pressure outlet (poiseuille/inputs3d_inflow), with the thresholds those
    inputs set (test.max_rel_err, test.max_transverse, test.max_rho_spread),
    and the Poiseuille grid convergence of poiseuille-regression.yml:98 in
    2-D and 3-D.
The 3-D walled runs take the two-step pair kernel with walls in FP64."""
import os

import numpy as np
import pytest

import anchors as A
from muphfasa_amd import host

pytestmark = pytest.mark.gpu

INP = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inputs")


def _read(name):
    return open(os.path.join(INP, name)).read()


def _couette3d(lattice):
    text = _read("couette_3d").replace("lbm.lattice_type = D3Q19", f"lbm.lattice_type = {lattice}")
    assert f"= {lattice}" in text
    return text


def test_couette_startup_2d():
    err, rate_err = A.couette_startup(A.host_snapshots, _read("couette_2d"))
    print("couette 2d", err, rate_err, flush=True)
    assert err <= 0.01 and rate_err <= 0.01


@pytest.mark.parametrize("lattice", ["D3Q15", "D3Q19", "D3Q27"])
def test_couette_startup_3d(lattice):
    err, rate_err = A.couette_startup(A.host_snapshots, _couette3d(lattice))
    print("couette 3d", lattice, err, rate_err, flush=True)
    assert err <= 0.01 and rate_err <= 0.01


def test_couette_convergence_2d():
    errs = A.couette_convergence(A.host_snapshots, _read("couette_2d"), [16, 32, 64])
    orders = A.observed_orders(errs)
    print("couette 2d convergence", errs, orders, flush=True)
    assert min(orders) >= 1.8


@pytest.mark.parametrize("lattice", ["D3Q15", "D3Q19", "D3Q27"])
def test_couette_convergence_3d(lattice):
    errs = A.couette_convergence(A.host_snapshots, _couette3d(lattice), [16, 32, 64])
    orders = A.observed_orders(errs)
    print("couette 3d convergence", lattice, errs, orders, flush=True)
    assert min(orders) >= 1.8


def test_tgv2d_convergence_and_decay():
    text = _read("tgv2d")
    errs = A.tgv2d_convergence(A.host_snapshots, text, [32, 64, 128])
    orders = A.observed_orders(errs)
    err, rate_err = A.tgv2d_decay(A.host_snapshots, text)
    print("tgv2d", errs, orders, err, rate_err, flush=True)
    assert min(orders) >= 1.8
    assert err <= 0.01 and rate_err <= 0.01


def test_tgv2d_start_is_the_analytic_field():
    """lbm.init = tgv2d sets the vortex of TgvAnalytic.H at t = 0 (u0 from
    units.ref_velocity): the first populations are the equilibrium of that
    field, so their own moments (unstreamed) give it back."""
    from oracle import oracle as O
    text = _read("tgv2d")
    u = host.inputs_units(text)
    sim = host.Simulation(text)
    sim.initialize()
    L = O.Lattice("D2Q9")
    fab = O.new_fab(L, [64, 64])
    O.valid(L, fab)[...] = sim.f()
    rho, vel = O.moments(L, fab, [64, 64], streamed=False)
    r0, u0, v0 = A.tgv2d(64, u["u_lat"], 0.0, 0.0)
    assert np.abs(vel[0, 0] - u0).max() < 1e-15 and np.abs(vel[1, 0] - v0).max() < 1e-15
    assert np.abs(rho[0] - r0).max() < 1e-14


@pytest.mark.parametrize("case,window,tol,trans,spread", [
    ("poiseuille_3d", (0.0, 1.0), 1.5e-3, 1e-6, 1e-8),
    ("poiseuille_3d_inflow", (0.5, 0.8), 5e-3, 4e-5, 2e-5),
])
def test_poiseuille_3d_regression(case, window, tol, trans, spread):
    """tests/regression/poiseuille/main.cpp:130-229 on inputs3d and
    inputs3d_inflow: profile and mean-velocity error against F H^2/(8 nu),
    transverse velocity and density spread over the window."""
    text = _read(case)
    u = host.inputs_units(text)
    max_step = A.step_of(text)
    (_, t, rho, vel), = list(A.host_snapshots(text, (), [max_step]))
    Nx, H = int(A.parse_inputs(text)["grid.ncells"][0]), float(A.parse_inputs(text)["grid.ncells"][2])
    nu = (u["tau"] - 0.5) / 3.0
    F = u["F0"]
    U_max = F * H * H / (8.0 * nu)
    i_lo, i_hi = int(round(window[0] * Nx)), int(round(window[1] * Nx))
    err, mean_u, tr, rs = A.poiseuille_window(rho, vel, 3, i_lo, i_hi, H, F, nu)
    mean_err = abs(mean_u - 2.0 / 3.0 * U_max) / U_max
    print(case, err / U_max, mean_err, tr / U_max, rs, flush=True)
    assert err / U_max < tol and mean_err < tol and tr / U_max < trans and rs < spread


@pytest.mark.parametrize("case,lattice", [("poiseuille_periodic", "D2Q9"), ("poiseuille_3d", "D3Q19"),
                                          ("poiseuille_3d", "D3Q27")])
def test_poiseuille_convergence(case, lattice):
    """poiseuille/convergence.py:37-60 with its defaults: wall-normal 16/32/64
    cells, 4 lateral cells, tau 0.8, run to nu t/H^2 = 2.5; inf-norm u_x error
    in physical units; observed order >= 1.8."""
    text = _read(case)
    D = 2 if lattice == "D2Q9" else 3
    nu = (0.8 - 0.5) / 3.0
    errs = []
    for n in (16, 32, 64):
        step = max(1, round(2.5 * n * n / nu))
        lat = " ".join(["4"] * (D - 1))
        lat_hi = " ".join([str(4 / n)] * (D - 1))
        ov = [f"grid.ncells={lat} {n}", f"geometry.prob_hi={lat_hi} 1.0", f"grid.max_grid_size={n}",
              "lbm.tau=0.8", f"sim.max_step={step}", f"lbm.lattice_type={lattice}"]
        u = host.inputs_units(text, ov)
        (_, t, rho, vel), = list(A.host_snapshots(text, ov, [step]))
        ua = A.poiseuille(float(n), u["F0"], nu)
        shape = (-1,) + (1,) * (D - 1)
        errs.append(float(np.abs(A.wall_field(vel, D) - ua.reshape(shape)).max()) * u["dx"] / u["dt"])
    orders = A.observed_orders(errs)
    print("poiseuille convergence", case, lattice, errs, orders, flush=True)
    assert min(orders) >= 1.8
