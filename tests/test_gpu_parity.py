"""GPU parity: every C-ABI entry point against the CPU oracle on the same
seeded inputs.  FP64 results are required to be identical to the oracle
(np.array_equal; NaN sentinels must sit at the same ghost positions): both
sides evaluate the reference's expressions in the same order without FMA."""
import zlib

import numpy as np
import pytest

from muphfasa_amd import abi
from oracle import oracle as O
from tests.gpu_util import (BC_SETS, bc_tables, ghosted, random_state, scalar_fields, to_dev,
                            to_host, valid_only)

pytestmark = pytest.mark.gpu

LATS = ["D2Q9", "D3Q15", "D3Q19", "D3Q27"]
SIZES = {2: [[17, 9], [64, 33], [1, 5]], 3: [[13, 7, 6], [5, 9, 11], [64, 3, 2]]}


def _shape_ids():
    out = []
    for lat in LATS:
        D = 2 if lat == "D2Q9" else 3
        for si, n in enumerate(SIZES[D]):
            for bc in BC_SETS:
                if n[0] == 1 and bc != "periodic":
                    continue
                out.append((lat, si, bc))
    return out


def _oracle_step(L, f, n, per, ref, omega, F):
    f = f.copy()
    O.fill_periodic(L, f, n, per)
    O.apply_bc(L, f, n, per, ref)
    out = O.new_fab(L, n)
    O.step(L, f, out, n, omega, F)
    return f, out


@pytest.mark.parametrize("lat,si,bc", _shape_ids())
@pytest.mark.parametrize("forced", [False, True])
def test_step_fused_and_halo_match_oracle(dev, lat, si, bc, forced):
    import torch
    L = O.Lattice(lat)
    n = SIZES[L.D][si]
    print(f"case lattice={lat} n={n} bc={bc} forced={forced}", flush=True)
    g, ref, per = bc_tables(L, bc)
    F = [2e-5, -1e-5, 3e-6][: L.D] if forced else None
    omega = 1.0 / 0.8
    f0 = random_state(L, n, seed=zlib.crc32(f"{lat}/{si}/{bc}".encode()))
    filled, expect = _oracle_step(L, f0, n, per, ref, omega, F)

    # fused: valid-only storage, wrap and wall values inline
    tin, fin = valid_only(L, n, dev, f0)
    tout, fout = valid_only(L, n, dev)
    abi.step(fin, fout, g, omega, F)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(tout), O.valid(L, expect))

    # halo: literal FillBoundary + PhysBCFunct + pullCollideFusedMF
    th, fh = ghosted(L, n, dev, f0)
    abi.fill_periodic(fh, g)
    abi.apply_bc(fh, g)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(th), filled, equal_nan=True)
    to2, fo2 = valid_only(L, n, dev)
    abi.pull_collide(fh, fo2, abi.LATTICE[lat], omega, F)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(to2), O.valid(L, expect))


@pytest.mark.parametrize("lat", LATS)
@pytest.mark.parametrize("forced", [False, True])
@pytest.mark.parametrize("streamed", [True, False])
def test_moments_match_oracle(dev, lat, forced, streamed):
    import torch
    L = O.Lattice(lat)
    n = SIZES[L.D][0]
    g, ref, per = bc_tables(L, "lid")
    F = [1e-4, 2e-5, -3e-5][: L.D] if forced else None
    f0 = random_state(L, n, seed=11)
    filled = f0.copy()
    O.fill_periodic(L, filled, n, per)
    O.apply_bc(L, filled, n, per, ref)
    rho_e, vel_e = O.moments(L, filled, n, F, streamed=streamed)
    tin, fin = valid_only(L, n, dev, f0)
    r, fr, v, fv = scalar_fields(L, n, dev)
    abi.moments(fin, fr, fv, g, F, streamed)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(r)[0], rho_e)
    assert np.array_equal(to_host(v), vel_e)


@pytest.mark.parametrize("lat", LATS)
@pytest.mark.parametrize("forced", [False, True])
def test_init_equilibrium_matches_oracle(dev, lat, forced):
    import torch
    L = O.Lattice(lat)
    n = SIZES[L.D][1]
    nn = O.n3(n)
    rng = np.random.default_rng(21)
    rho = 1.0 + 0.02 * rng.standard_normal(nn[::-1])
    vel = 0.03 * rng.standard_normal((L.D,) + tuple(nn[::-1]))
    F = [1e-3, -2e-4, 5e-4][: L.D] if forced else None
    f_e, rho_e, vel_e = O.init_equilibrium(L, rho, vel, n, F)
    r, fr, v, fv = scalar_fields(L, n, dev, rho, vel)
    tf, ff = valid_only(L, n, dev)
    abi.init_equilibrium(fr, fv, ff, abi.LATTICE[lat], F)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(tf), O.valid(L, f_e))
    assert np.array_equal(to_host(r)[0], rho_e)
    assert np.array_equal(to_host(v), vel_e)


@pytest.mark.parametrize("lat", LATS)
def test_reductions(dev, lat):
    import math
    import torch
    L = O.Lattice(lat)
    n = [37, 21, 9] if L.D == 3 else [129, 77]
    f0 = random_state(L, n, seed=5)
    tin, fin = valid_only(L, n, dev, f0)
    out = torch.zeros(32, dtype=torch.float64, device=dev)
    v = O.valid(L, f0)
    abi.reduce(fin, abi.LATTICE[lat], abi.RED["mass"], out.data_ptr())
    torch.cuda.synchronize()
    exact = math.fsum(v.ravel())
    assert abs(out[0].item() - exact) <= 1e-13 * abs(exact)
    abi.reduce(fin, abi.LATTICE[lat], abi.RED["momentum"], out.data_ptr())
    torch.cuda.synchronize()
    for d in range(L.D):
        exact = math.fsum((L.c[:, d][:, None, None, None] * v).ravel())
        assert abs(out[d].item() - exact) <= 1e-13 * math.fsum(np.abs(v).ravel())
    abi.reduce(fin, abi.LATTICE[lat], abi.RED["comp_sum"], out.data_ptr())
    torch.cuda.synchronize()
    for q in range(L.Q):
        exact = math.fsum(v[q].ravel())
        assert abs(out[q].item() - exact) <= 1e-13 * abs(exact)
    # deterministic: same input, same bits
    abi.reduce(fin, abi.LATTICE[lat], abi.RED["mass"], out.data_ptr())
    a = out[0].item()
    abi.reduce(fin, abi.LATTICE[lat], abi.RED["mass"], out.data_ptr())
    torch.cuda.synchronize()
    assert out[0].item() == a
    # residual
    t2 = tin.clone()
    t2[0, 0, 0, 0] += 0.5
    f2 = abi.field(t2, n, 0, L.D)
    abi.residual(f2, fin, out.data_ptr())
    torch.cuda.synchronize()
    assert abs(out[0].item() - 0.25) < 1e-15
    assert abs(out[1].item() - math.fsum((to_host(t2) ** 2).ravel())) < 1e-12 * out[1].item()


def test_empty_and_mismatched_boxes_are_errors(dev):
    import torch
    L = O.Lattice("D3Q19")
    g, ref, per = bc_tables(L, "periodic")
    a = torch.zeros((19, 4, 4, 4), dtype=torch.float64, device=dev)
    b = torch.zeros((19, 4, 4, 5), dtype=torch.float64, device=dev)
    fa = abi.field(a, [4, 4, 4], 0, 3)
    fb = abi.field(b, [5, 4, 4], 0, 3)
    with pytest.raises(abi.MfaError):
        abi.step(fa, fb, g, 1.0)
    with pytest.raises(abi.MfaError):
        abi.step(fa, fa, g, 1.0)         # in-place is rejected (pull needs two lattices)
    fz = abi.field(a, [4, 4, 4], 0, 3)
    fz.n[0] = 0
    with pytest.raises(abi.MfaError):
        abi.step(fz, fz, g, 1.0)


LAUNCHES = [(256, 1, 1, 0), (128, 2, 2, 1), (64, 4, 2, 2), (64, 1, 1, 2), (128, 1, 2, 0)]


@pytest.mark.parametrize("lat", LATS)
@pytest.mark.parametrize("bc", ["periodic", "lid", "inflow_pressure"])
def test_every_launch_configuration_matches_oracle(dev, lat, bc):
    """Block shapes, two cells per thread along z (odd n_z leaves a partial
    column) and the three load policies all give the oracle's bits."""
    import torch
    L = O.Lattice(lat)
    n = [70, 9, 7] if L.D == 3 else [70, 13]
    g, ref, per = bc_tables(L, bc)
    F = [2e-5, -1e-5, 3e-6][: L.D]
    f0 = random_state(L, n, seed=zlib.crc32(f"launch/{lat}/{bc}".encode()))
    _, expect = _oracle_step(L, f0, n, per, ref, 1.0 / 0.8, F)
    lid = abi.LATTICE[lat]
    tin, fin = valid_only(L, n, dev, f0)
    try:
        for cfg in LAUNCHES:                # (2-D runs one cell per thread whatever cfg says)
            abi.set_launch(lid, *cfg)
            tout, fout = valid_only(L, n, dev)
            abi.step(fin, fout, g, 1.0 / 0.8, F)
            torch.cuda.synchronize()
            assert np.array_equal(to_host(tout), O.valid(L, expect)), cfg
    finally:
        abi.set_launch(lid)


def _to_rows(fab):
    """(Q, e2, e1, e0) component-major -> (e2, e1, Q, e0) row-interleaved."""
    return np.ascontiguousarray(np.transpose(fab, (1, 2, 0, 3)))


def _from_rows(a):
    return np.ascontiguousarray(np.transpose(a, (2, 0, 1, 3)))


@pytest.mark.parametrize("lat", LATS)
@pytest.mark.parametrize("bc", ["periodic", "lid", "inflow_pressure", "xper_walls"])
def test_row_interleaved_layout_matches_oracle(dev, lat, bc):
    """The row-interleaved population order (the Q rows of one (y, z) side by
    side) runs the same kernels through strides: the fused step under every
    launch configuration, the literal ghost-fill + pull-collide path and the
    moments all give the oracle's bits."""
    import torch
    L = O.Lattice(lat)
    n = [70, 9, 7] if L.D == 3 else [70, 13]
    print(f"case row layout lattice={lat} bc={bc}", flush=True)
    g, ref, per = bc_tables(L, bc)
    F = [2e-5, -1e-5, 3e-6][: L.D]
    omega = 1.0 / 0.8
    f0 = random_state(L, n, seed=zlib.crc32(f"rows/{lat}/{bc}".encode()))
    filled, expect = _oracle_step(L, f0, n, per, ref, omega, F)
    lid = abi.LATTICE[lat]
    tin = to_dev(_to_rows(O.valid(L, f0)), dev)
    fin = abi.field(tin, n, 0, L.D, layout="row")
    assert fin.sq == n[0] and fin.sy == L.Q * n[0]
    try:
        for cfg in LAUNCHES:
            abi.set_launch(lid, *cfg)
            tout = torch.full_like(tin, float("nan"))
            abi.step(fin, abi.field(tout, n, 0, L.D, layout="row"), g, omega, F)
            torch.cuda.synchronize()
            assert np.array_equal(_from_rows(to_host(tout)), O.valid(L, expect)), cfg
    finally:
        abi.set_launch(lid)
    th = to_dev(_to_rows(f0), dev)
    fh = abi.field(th, n, 1, L.D, layout="row")
    abi.fill_periodic(fh, g)
    abi.apply_bc(fh, g)
    torch.cuda.synchronize()
    assert np.array_equal(_from_rows(to_host(th)), filled, equal_nan=True)
    to2 = torch.full_like(tin, float("nan"))
    abi.pull_collide(fh, abi.field(to2, n, 0, L.D, layout="row"), lid, omega, F)
    torch.cuda.synchronize()
    assert np.array_equal(_from_rows(to_host(to2)), O.valid(L, expect))
    rho_e, vel_e = O.moments(L, filled, n, F, streamed=True)
    r, fr, v, fv = scalar_fields(L, n, dev)
    abi.moments(fin, fr, fv, g, F, True)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(r)[0], rho_e)
    assert np.array_equal(to_host(v), vel_e)


STEP2_SIZES = [[13, 7, 6], [64, 3, 2], [40, 33, 70], [1, 5, 3], [70, 20, 150]]


@pytest.mark.parametrize("lat", ["D3Q15", "D3Q19", "D3Q27"])
@pytest.mark.parametrize("si", range(len(STEP2_SIZES)))
@pytest.mark.parametrize("forced", [False, True])
@pytest.mark.parametrize("layout", ["component", "row"])
@pytest.mark.parametrize("bc", ["periodic", "noslip", "lid", "xper_walls", "inflow_pressure",
                                "inflow_outflow"])
def test_step2_equals_two_oracle_steps(dev, lat, si, forced, layout, bc):
    """mfa_step2 (two steps per pass, intermediate state in LDS) gives the
    oracle's bits after two steps: boxes smaller than its tile, not a
    multiple of it, one cell wide, and taller than one block's z-range; walls
    of every kind on any face (the second step's ghost values come from the
    intermediate state)."""
    import torch
    L = O.Lattice(lat)
    n = STEP2_SIZES[si]
    print(f"case step2 lattice={lat} n={n} forced={forced} layout={layout} bc={bc}", flush=True)
    g, ref, per = bc_tables(L, bc)
    F = [2e-5, -1e-5, 3e-6] if forced else None
    omega = 1.0 / 0.8
    f0 = random_state(L, n, seed=zlib.crc32(f"step2/{lat}/{si}/{bc}".encode()))
    _, mid = _oracle_step(L, f0, n, per, ref, omega, F)
    _, expect = _oracle_step(L, mid, n, per, ref, omega, F)
    if layout == "row":
        tin = to_dev(_to_rows(O.valid(L, f0)), dev)
    else:
        tin = to_dev(np.ascontiguousarray(O.valid(L, f0)), dev)
    tout = torch.full_like(tin, float("nan"))
    abi.step2(abi.field(tin, n, 0, 3, layout=layout), abi.field(tout, n, 0, 3, layout=layout),
              g, omega, F)
    torch.cuda.synchronize()
    got = _from_rows(to_host(tout)) if layout == "row" else to_host(tout)
    assert np.array_equal(got, O.valid(L, expect))


@pytest.mark.parametrize("lat", ["D3Q19", "D3Q27"])
@pytest.mark.parametrize("layout", ["component", "row"])
@pytest.mark.parametrize("dt,bc", [("f64", "periodic"), ("f64", "lid"), ("f32", "periodic")])
def test_step2_writes_only_the_box(dev, lat, layout, dt, bc):
    """mfa_step2 on a box held inside a larger buffer (a ghost layer plus
    extra padding on every side; the box cuts the 64-cell columns in x and
    the tile rows in y): the box gets the bits of two steps (the oracle's for
    FP64, two mfa_step calls' for FP32 storage) and every other word of the
    output buffer keeps its sentinel.  The pair kernels drop the stores of
    lanes outside the box in the hardware (an offset beyond the store's
    range) instead of branching around them; this pins that."""
    import torch
    L = O.Lattice(lat)
    lid, did = abi.LATTICE[lat], abi.DTYPE[dt]
    n = [70, 13, 9]
    print(f"case step2_box lattice={lat} layout={layout} dtype={dt} bc={bc}", flush=True)
    g, ref, per = bc_tables(L, bc)
    F = [2e-5, -1e-5, 3e-6]
    omega = 1.0 / 0.8
    f0 = random_state(L, n, seed=zlib.crc32(f"step2_box/{lat}".encode()))
    e0, e1, e2 = n[0] + 2 + 5, n[1] + 2 + 3, n[2] + 2 + 1
    tdt = torch.float64 if dt == "f64" else torch.float32
    if layout == "component":
        shape = (L.Q, e2, e1, e0)
        box = (slice(None), slice(1, 1 + n[2]), slice(1, 1 + n[1]), slice(1, 1 + n[0]))
    else:
        shape = (e2, e1, L.Q, e0)
        box = (slice(1, 1 + n[2]), slice(1, 1 + n[1]), slice(None), slice(1, 1 + n[0]))
    sentinel = -12345.5
    outside = torch.ones(shape, dtype=torch.bool, device=dev)
    outside[box] = False

    def padded():
        return torch.full(shape, sentinel, dtype=tdt, device=dev)

    def fld(t):
        return abi.field(t, n, 1, 3, did, layout=layout)

    def in_layout(comp):                       # (Q, z, y, x) -> the buffer's order
        return comp if layout == "component" else comp.permute(1, 2, 0, 3)

    src = to_dev(np.ascontiguousarray(O.valid(L, f0)), dev)
    fin = padded()
    if dt == "f64":
        fin[box] = in_layout(src)
    else:
        dev32 = torch.empty((L.Q, n[2], n[1], n[0]), dtype=tdt, device=dev)
        abi.convert(abi.field(src, n, 0, 3), abi.field(dev32, n, 0, 3, did), lid)
        fin[box] = in_layout(dev32)
    fout = padded()
    abi.step2(fld(fin), fld(fout), g, omega, F)
    if dt == "f64":
        _, mid = _oracle_step(L, f0, n, per, ref, omega, F)
        _, expect = _oracle_step(L, mid, n, per, ref, omega, F)
        want = in_layout(to_dev(np.ascontiguousarray(O.valid(L, expect)), dev))
    else:
        mid, two = padded(), padded()
        abi.step(fld(fin), fld(mid), g, omega, F)
        abi.step(fld(mid), fld(two), g, omega, F)
        want = two[box]
    torch.cuda.synchronize()
    word = torch.int64 if dt == "f64" else torch.int32
    assert torch.equal(fout[box].contiguous().view(word), want.contiguous().view(word))
    assert bool((fout[outside] == sentinel).all())


STEP2_2D_SIZES = [[13, 7], [64, 3], [130, 9], [1, 5], [70, 33]]


@pytest.mark.parametrize("bc", ["periodic", "noslip", "lid", "xper_walls", "inflow_pressure", "inflow_outflow"])
@pytest.mark.parametrize("si", range(len(STEP2_2D_SIZES)))
@pytest.mark.parametrize("forced", [False, True])
def test_step2_2d_equals_two_oracle_steps(dev, bc, si, forced):
    """mfa_step2 on D2Q9 boxes, periodic or with walls of every kind (the
    second step's ghost values come from the intermediate state in LDS):
    the oracle's bits after two steps, for boxes smaller than the 64 x 4
    tile, not a multiple of it, and one cell wide, in both population
    orders."""
    L = O.Lattice("D2Q9")
    n = STEP2_2D_SIZES[si]
    print(f"case step2_2d bc={bc} n={n} forced={forced}", flush=True)
    g, ref, per = bc_tables(L, bc)
    F = [2e-5, -1e-5] if forced else None
    omega = 1.0 / 0.7
    f0 = random_state(L, n, seed=zlib.crc32(f"step2_2d/{bc}/{si}".encode()))
    _, mid = _oracle_step(L, f0, n, per, ref, omega, F)
    _, expect = _oracle_step(L, mid, n, per, ref, omega, F)
    import torch
    for layout in ("component", "row"):
        if layout == "row":
            tin = to_dev(_to_rows(O.valid(L, f0)), dev)
        else:
            tin = to_dev(np.ascontiguousarray(O.valid(L, f0)), dev)
        tout = torch.full_like(tin, float("nan"))
        abi.step2(abi.field(tin, n, 0, 2, layout=layout), abi.field(tout, n, 0, 2, layout=layout),
                  g, omega, F)
        torch.cuda.synchronize()
        got = _from_rows(to_host(tout)) if layout == "row" else to_host(tout)
        assert np.array_equal(got, O.valid(L, expect)), layout


def test_step2_refuses_3d_walls_in_fp32_and_2d_16bit(dev):
    import torch
    L = O.Lattice("D3Q19")
    g, _, _ = bc_tables(L, "lid")
    a = torch.zeros((19, 4, 4, 4), dtype=torch.float32, device=dev)
    b = torch.zeros_like(a)
    with pytest.raises(abi.MfaError) as e:
        abi.step2(abi.field(a, [4, 4, 4], 0, 3, abi.DTYPE["f32"]),
                  abi.field(b, [4, 4, 4], 0, 3, abi.DTYPE["f32"]), g, 1.0)
    assert e.value.code == -2
    L2 = O.Lattice("D2Q9")
    g2, _, _ = bc_tables(L2, "periodic")
    a2 = torch.zeros((9, 1, 4, 4), dtype=torch.float16, device=dev)
    with pytest.raises(abi.MfaError) as e:
        abi.step2(abi.field(a2, [4, 4], 0, 2, abi.DTYPE["f16dev"]),
                  abi.field(a2.clone(), [4, 4], 0, 2, abi.DTYPE["f16dev"]), g2, 1.0)
    assert e.value.code == -2
