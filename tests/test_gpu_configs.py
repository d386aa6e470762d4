"""Multi-step parity on the benchmark configurations.

* D2Q9 lid cavity 128x128, Re=100 (tau=0.884), 10 000 steps: every
  population, rho and u identical to the oracle run of the same steps.
* D3Q19 periodic box with Guo force, seeded noise: identical to the oracle
  at 32^3 x 200 steps; at the full 256^3 the size-independent invariants are
  checked instead (FP64 mass conserved, momentum grows by exactly N^3 F per
  step, to rounding).
"""
import zlib

import numpy as np
import pytest

from muphfasa_amd import abi
from oracle import oracle as O
from tests.gpu_util import bc_tables, scalar_fields, to_dev, to_host, valid_only

pytestmark = pytest.mark.gpu


def _run_gpu(g, ta, fa, tb, fb, omega, F, nsteps):
    for _ in range(nsteps):
        abi.step(fa, fb, g, omega, F)
        ta, fa, tb, fb = tb, fb, ta, fa
    return ta, fa


def test_lid_cavity_128_10000_steps(dev):
    import torch
    L = O.Lattice("D2Q9")
    N = 128
    n = [N, N]
    u_lid = 0.1
    nu = 0.128                       # Re=100, u_lat=0.1, dx=1/128 -> tau=0.884
    tau = 0.5 + nu * 3.0
    omega = 1.0 / tau
    g, ref, per = bc_tables(L, "lid")
    # lid velocity of this config: (0.1, 0)
    fu = np.zeros((6, 3))
    fu[3, 0] = u_lid
    ft = [abi.BC["NoSlip"]] * 3 + [abi.BC["MovingWall"]]
    g = abi.resolve_bc(0, ft, [1.0] * 6, fu)
    ref = O.resolve_bc(2, ft, np.ones(6), fu)
    rho0 = np.ones((1, N, N))
    vel0 = np.zeros((2, 1, N, N))
    f0, _, _ = O.init_equilibrium(L, rho0, vel0, n)
    O.apply_bc(L, f0, n, [0, 0, 0], ref)
    print("oracle: 10000 steps of 128^2 D2Q9", flush=True)
    fe = O.run(L, f0, n, omega, None, [0, 0], ref, 10000)
    print("gpu: 10000 steps", flush=True)
    ta, fa = valid_only(L, n, dev, f0)
    tb, fb = valid_only(L, n, dev)
    ta, fa = _run_gpu(g, ta, fa, tb, fb, omega, None, 10000)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(ta), O.valid(L, fe))
    rho_e, vel_e = O.moments(L, fe, n)
    r, fr, v, fv = scalar_fields(L, n, dev)
    abi.moments(fa, fr, fv, g, None, True)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(r)[0], rho_e)
    assert np.array_equal(to_host(v), vel_e)
    assert np.abs(vel_e).max() > 0.01


def _periodic_box_ic(n, seed):
    rng = np.random.default_rng(seed)
    nn = O.n3(n)
    rho = np.ones(nn[::-1])
    vel = rng.uniform(-1e-3, 1e-3, (3,) + tuple(nn[::-1]))
    return rho, vel


def test_periodic_forced_d3q19_32cubed_parity(dev):
    import torch
    L = O.Lattice("D3Q19")
    n = [32, 32, 32]
    F = [1e-6, 0.0, 0.0]
    omega = 1.0 / 0.8
    g, ref, per = bc_tables(L, "periodic")
    rho, vel = _periodic_box_ic(n, 2024)
    f0, _, _ = O.init_equilibrium(L, rho, vel, n, F)
    O.fill_periodic(L, f0, n, [1, 1, 1])
    fe = O.run(L, f0, n, omega, F, [1, 1, 1], ref, 200)
    r, fr, v, fv = scalar_fields(L, n, dev, rho, vel)
    ta, fa = valid_only(L, n, dev)
    tb, fb = valid_only(L, n, dev)
    abi.init_equilibrium(fr, fv, fa, 2, F)
    ta, fa = _run_gpu(g, ta, fa, tb, fb, omega, F, 200)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(ta), O.valid(L, fe))


def test_periodic_forced_d3q19_256cubed_invariants(dev):
    """The bench workload at full size in both population orders: FP64 mass
    conserved, x momentum grows by exactly N^3 F per step, and the
    row-interleaved run (the bench's order) is bit-identical to the
    component-major one after the same steps."""
    import torch
    N = 256
    n = [N, N, N]
    F = [1e-6, 0.0, 0.0]
    omega = 1.0 / 0.8
    g = abi.resolve_bc(2, [abi.BC["Periodic"]] * 6)
    gen = torch.Generator(device=dev)
    gen.manual_seed(7)
    rho = torch.ones((1, N, N, N), dtype=torch.float64, device=dev)
    vel = (torch.rand((3, N, N, N), dtype=torch.float64, device=dev, generator=gen) * 2 - 1) * 1e-3
    fr, fv = abi.field(rho, n, 0, 3), abi.field(vel, n, 0, 3)
    steps = 50
    final = {}
    for layout in ("component", "row"):
        print(f"case d3q19 256^3 layout={layout}", flush=True)
        shape = (19, N, N, N) if layout == "component" else abi.row_shape(19, n, 0, 3)
        ta = torch.empty(shape, dtype=torch.float64, device=dev)
        tb = torch.empty_like(ta)
        fa, fb = abi.field(ta, n, 0, 3, layout=layout), abi.field(tb, n, 0, 3, layout=layout)
        r0, v0 = rho.clone(), vel.clone()
        abi.init_equilibrium(abi.field(r0, n, 0, 3), abi.field(v0, n, 0, 3), fa, 2, F)
        out = torch.zeros(8, dtype=torch.float64, device=dev)
        abi.reduce(fa, 2, abi.RED["mass"], out.data_ptr())
        abi.reduce(fa, 2, abi.RED["momentum"], out[1:].data_ptr())
        torch.cuda.synchronize()
        m0, p0 = out[0].item(), out[1:4].cpu().numpy().copy()
        ta, fa = _run_gpu(g, ta, fa, tb, fb, omega, F, steps)
        abi.reduce(fa, 2, abi.RED["mass"], out.data_ptr())
        abi.reduce(fa, 2, abi.RED["momentum"], out[1:].data_ptr())
        torch.cuda.synchronize()
        m1, p1 = out[0].item(), out[1:4].cpu().numpy()
        assert abs(m1 - m0) / m0 < 1e-13
        gain = steps * N ** 3 * F[0]
        assert abs((p1[0] - p0[0]) - gain) / gain < 1e-9
        assert abs(p1[1] - p0[1]) < 1e-9 and abs(p1[2] - p0[2]) < 1e-9
        assert torch.isfinite(ta).all().item()
        final[layout] = ta if layout == "component" else ta.permute(2, 0, 1, 3)
        del tb
    assert torch.equal(final["component"], final["row"])


def test_step2_equals_two_steps_at_256cubed(dev):
    """The bench workload: one mfa_step2 pass equals two mfa_step calls bit
    for bit at the full size, in the row-interleaved order."""
    import torch
    N = 256
    n = [N, N, N]
    F = [1e-6, 0.0, 0.0]
    omega = 1.0 / 0.8
    g = abi.resolve_bc(2, [abi.BC["Periodic"]] * 6)
    gen = torch.Generator(device=dev)
    gen.manual_seed(11)
    rho = torch.ones((1, N, N, N), dtype=torch.float64, device=dev)
    vel = (torch.rand((3, N, N, N), dtype=torch.float64, device=dev, generator=gen) * 2 - 1) * 1e-3
    shape = abi.row_shape(19, n, 0, 3)
    ta = torch.empty(shape, dtype=torch.float64, device=dev)
    fa = abi.field(ta, n, 0, 3, layout="row")
    abi.init_equilibrium(abi.field(rho, n, 0, 3), abi.field(vel, n, 0, 3), fa, 2, F)
    del rho, vel
    tb = torch.empty_like(ta)
    tc = torch.empty_like(ta)
    fb, fc = abi.field(tb, n, 0, 3, layout="row"), abi.field(tc, n, 0, 3, layout="row")
    abi.step(fa, fb, g, omega, F)
    abi.step(fb, fc, g, omega, F)
    abi.step2(fa, fb, g, omega, F)
    torch.cuda.synchronize()
    assert torch.equal(tb, tc)


@pytest.mark.parametrize("dt,side", [("f64", 5300), ("f32", 7500)])
def test_step2_at_the_largest_plane_it_accepts(dev, dt, side):
    """mfa_step2 on a D3Q19 box (row order) whose z-plane spans just under the
    4 GiB less 16 bytes it accepts: side x side x 3 cells, 4.27e9 bytes per
    plane in FP64 (5300^2) and in FP32 storage (7500^2), rows and columns cut
    by the tiles.  The pair's 32-bit byte offsets reach the top of their range
    and stay below the offset of the dropped stores; the pass equals two
    mfa_step calls bit for bit."""
    import torch
    n = [side, side, 3]
    span = (n[0] + (n[1] - 1) * 19 * n[0]) * (8 if dt == "f64" else 4)
    assert 0xF0000000 < span <= 0xFFFFFFF0
    print(f"case step2 largest plane dtype={dt} n={n} span={span}", flush=True)
    did = abi.DTYPE[dt]
    tdt = torch.float64 if dt == "f64" else torch.float32
    F = [1e-6, 0.0, 0.0]
    omega = 1.0 / 0.8
    g = abi.resolve_bc(2, [abi.BC["Periodic"]] * 6)
    gen = torch.Generator(device=dev)
    gen.manual_seed(5)
    shape = abi.row_shape(19, n, 0, 3)
    ta = torch.empty(shape, dtype=tdt, device=dev)
    fa = abi.field(ta, n, 0, 3, did, layout="row")
    rho = torch.ones((1, n[2], n[1], n[0]), dtype=torch.float64, device=dev)
    vel = (torch.rand((3, n[2], n[1], n[0]), dtype=torch.float64, device=dev, generator=gen) * 2 - 1) * 1e-2
    abi.init_equilibrium(abi.field(rho, n, 0, 3), abi.field(vel, n, 0, 3), fa, 2, F)
    del rho, vel
    tb = torch.empty_like(ta)
    tc = torch.empty_like(ta)
    fb = abi.field(tb, n, 0, 3, did, layout="row")
    fc = abi.field(tc, n, 0, 3, did, layout="row")
    abi.step(fa, fb, g, omega, F)
    abi.step(fb, fc, g, omega, F)
    tb.fill_(float("nan"))
    abi.step2(fa, fb, g, omega, F)
    torch.cuda.synchronize()
    word = torch.int64 if dt == "f64" else torch.int32
    same = torch.equal(tb.view(word), tc.view(word))
    del ta, tb, tc
    torch.cuda.empty_cache()
    assert same


@pytest.mark.parametrize("lat", ["D3Q15", "D3Q19", "D3Q27"])
@pytest.mark.parametrize("bc", ["lid", "xper_walls", "inflow_pressure"])
def test_walled_step2_equals_two_steps_at_config_scale(dev, lat, bc):
    """Walled 3-D boxes take their pairs as a periodic pass plus the wall
    shell (two launches).  On a box of many tiles in every direction (not a
    multiple of the pair kernel's 64 x 8 columns nor of the shell kernel's
    16 x 8 face tiles), with a forced, non-uniform state, one mfa_step2 equals
    two mfa_step calls bit for bit, in both population orders."""
    import torch
    L = O.Lattice(lat)
    n = [150, 97, 83]
    print(f"case walled step2 lattice={lat} bc={bc} n={n}", flush=True)
    g, _, _ = bc_tables(L, bc)
    lid = abi.LATTICE[lat]
    F = [2e-6, -1e-6, 5e-7]
    omega = 1.0 / 0.62
    gen = torch.Generator(device=dev)
    gen.manual_seed(zlib.crc32(f"walled/{lat}/{bc}".encode()))
    rho = 1.0 + (torch.rand((1, n[2], n[1], n[0]), dtype=torch.float64, device=dev, generator=gen) - 0.5) * 2e-2
    vel = (torch.rand((3, n[2], n[1], n[0]), dtype=torch.float64, device=dev, generator=gen) * 2 - 1) * 3e-2
    for layout in ("row", "component"):
        shape = abi.row_shape(L.Q, n, 0, 3) if layout == "row" else (L.Q, n[2], n[1], n[0])
        ta = torch.empty(shape, dtype=torch.float64, device=dev)
        fa = abi.field(ta, n, 0, 3, layout=layout)
        abi.init_equilibrium(abi.field(rho, n, 0, 3), abi.field(vel, n, 0, 3), fa, lid, F)
        tb, tc, td = (torch.full_like(ta, float("nan")) for _ in range(3))
        fb, fc, fd = (abi.field(t, n, 0, 3, layout=layout) for t in (tb, tc, td))
        abi.step(fa, fb, g, omega, F)
        abi.step(fb, fc, g, omega, F)
        abi.step2(fa, fd, g, omega, F)
        torch.cuda.synchronize()
        assert torch.isfinite(tc).all().item()
        assert torch.equal(tc, td), layout
        del ta, tb, tc, td
