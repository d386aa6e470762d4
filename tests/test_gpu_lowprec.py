"""Reduced-precision storage (FP32 / FP16 / BF16 shifted deviations, FP32
collision).

* One step against a float32 reference of the deviation-form BGK+Guo update
  restated in numpy in tests/lowprec_ref.py (same inputs: the dequantised stored deviations),
  to within a few units of the storage type's last place.
* Round trip FP64 -> reduced -> FP64 within half an ulp of the storage type.
* Multi-step error against the FP64 engine on a 3-D Taylor-Green vortex with
  a stated bound per storage type, and FP64 mass bookkeeping.
"""
import numpy as np
import pytest

from muphfasa_amd import abi
from oracle import oracle as O
from tests.lowprec_ref import ref_step_f32 as _ref_step_f32

pytestmark = pytest.mark.gpu

TORCH_DT = {"f32": "float32", "f16dev": "float16", "bf16dev": "bfloat16"}
# unit roundoff of each storage type (relative to the stored deviation)
EPS = {"f32": 2.0 ** -24, "f16dev": 2.0 ** -11, "bf16dev": 2.0 ** -8}
# absolute rounding floor: half the spacing of IEEE half's subnormals (2^-24)
FLOOR = {"f32": 0.0, "f16dev": 2.0 ** -25, "bf16dev": 0.0}


def _dev_tensor(torch, Q, n, dt, dev):
    return torch.empty((Q, n[2], n[1], n[0]), dtype=getattr(torch, TORCH_DT[dt]), device=dev)


@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19", "D3Q27"])
@pytest.mark.parametrize("dt", ["f32", "f16dev", "bf16dev"])
@pytest.mark.parametrize("forced", [False, True])
def test_one_step_matches_float32_reference(dev, lat, dt, forced):
    import torch
    L = O.Lattice(lat)
    n = [24, 10, 7] if L.D == 3 else [40, 17, 1]
    rng = np.random.default_rng(3)
    f64 = L.w[:, None, None, None] * (1 + 0.05 * rng.standard_normal((L.Q, n[2], n[1], n[0])))
    src = torch.from_numpy(f64).to(dev)
    a = _dev_tensor(torch, L.Q, n, dt, dev)
    b = torch.empty_like(a)
    D = L.D
    fa = abi.field(a, n, 0, D, abi.DTYPE[dt])
    fb = abi.field(b, n, 0, D, abi.DTYPE[dt])
    abi.convert(abi.field(src, n, 0, D), fa, abi.LATTICE[lat])
    g = abi.resolve_bc(abi.LATTICE[lat], [abi.BC["Periodic"]] * 6)
    F = [2e-4, -1e-4, 5e-5][:D] if forced else None
    abi.step(fa, fb, g, 1.0 / 0.6, F)
    torch.cuda.synchronize()
    d_in = a.float().cpu().numpy()
    expect = _ref_step_f32(L, d_in, 1.0 / 0.6, F)
    got = b.float().cpu().numpy()
    # result rounded once to the storage type: |got - expect| <= ~1 ulp(T) of
    # the deviation plus float32 evaluation-order noise
    bound = (2.5 * EPS[dt] + 4 * 2.0 ** -24) * (np.abs(expect) + 1e-3 * L.w[:, None, None, None]) \
        + FLOOR[dt]
    excess = (np.abs(got - expect) / bound).max()
    assert excess <= 1.0, (dt, excess)


@pytest.mark.parametrize("dt", ["f32", "f16dev", "bf16dev"])
def test_convert_round_trip(dev, dt):
    import torch
    L = O.Lattice("D3Q19")
    n = [16, 8, 4]
    rng = np.random.default_rng(9)
    f64 = L.w[:, None, None, None] * (1 + 0.02 * rng.standard_normal((L.Q, n[2], n[1], n[0])))
    src = torch.from_numpy(f64).to(dev)
    a = _dev_tensor(torch, L.Q, n, dt, dev)
    back = torch.empty_like(src)
    abi.convert(abi.field(src, n, 0, 3), abi.field(a, n, 0, 3, abi.DTYPE[dt]), 2)
    abi.convert(abi.field(a, n, 0, 3, abi.DTYPE[dt]), abi.field(back, n, 0, 3), 2)
    torch.cuda.synchronize()
    dev_true = f64 - L.w[:, None, None, None]
    err = np.abs(back.cpu().numpy() - f64)
    assert np.all(err <= EPS[dt] * np.abs(dev_true) * 1.0001 + FLOOR[dt] + 1e-300)
    with pytest.raises(abi.MfaError):
        abi.convert(abi.field(src, n, 0, 3), abi.field(back, n, 0, 3), 2)


# bound on max |u - u_fp64| / u0 after the TGV run below, per storage type
TGV_BOUND = {"f32": 1e-4, "f16dev": 2e-2, "bf16dev": 1.5e-1}


@pytest.mark.parametrize("dt", ["f32", "f16dev", "bf16dev"])
def test_tgv_error_against_fp64(dev, dt):
    import torch
    L = O.Lattice("D3Q27")
    N = 48
    n = [N, N, N]
    u0 = 0.05
    tau = 0.6
    x = 2 * np.pi * (np.arange(N) + 0.5) / N
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    vel = np.stack([u0 * np.sin(X) * np.cos(Y) * np.cos(Z), -u0 * np.cos(X) * np.sin(Y) * np.cos(Z),
                    np.zeros_like(X)])
    rho = 1 + 3 * u0 ** 2 / 16 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2)
    g = abi.resolve_bc(3, [abi.BC["Periodic"]] * 6)
    out = {}
    for st in ("f64", dt):
        r = torch.from_numpy(rho[None].copy()).to(dev)
        v = torch.from_numpy(vel.copy()).to(dev)
        tdt = torch.float64 if st == "f64" else getattr(torch, TORCH_DT[st])
        a = torch.empty((27, N, N, N), dtype=tdt, device=dev)
        b = torch.empty_like(a)
        fa, fb = abi.field(a, n, 0, 3, abi.DTYPE[st]), abi.field(b, n, 0, 3, abi.DTYPE[st])
        abi.init_equilibrium(abi.field(r, n, 0, 3), abi.field(v, n, 0, 3), fa, 3)
        mt = torch.zeros(4, dtype=torch.float64, device=dev)
        abi.reduce(fa, 3, abi.RED["mass"], mt.data_ptr())
        for _ in range(300):
            abi.step(fa, fb, g, 1.0 / tau)
            fa, fb = fb, fa
        abi.reduce(fa, 3, abi.RED["mass"], mt[1:].data_ptr())
        abi.moments(fa, abi.field(r, n, 0, 3), abi.field(v, n, 0, 3), g)
        torch.cuda.synchronize()
        out[st] = (v.cpu().numpy(), mt[0].item(), mt[1].item())
    v64, m0, m1 = out["f64"]
    vlp, l0, l1 = out[dt]
    assert abs(m1 - m0) / m0 < 1e-13
    err = np.abs(vlp - v64).max() / u0
    print(f"{dt}: max |u - u_fp64| / u0 = {err:.3e}; mass drift {abs(l1 - l0) / l0:.3e}")
    assert err < TGV_BOUND[dt]
    assert np.isfinite(vlp).all()


PAIR_CASES = {
    # name: (n for 3-D, n for 2-D, boundary set, forced, launches to compare)
    "periodic_forced": ([40, 10, 7], [40, 17, 1], "periodic", True, [(64, 4, 1), (128, 2, 0)]),
    "lid_walls": ([24, 10, 7], [24, 17, 1], "lid", False, [(64, 4, 1)]),
    "multiwave_periodic": ([260, 6, 3], [260, 9, 1], "periodic", False, [(64, 4, 1), (256, 1, 1)]),
    "multiwave_walls": ([258, 5, 3], [258, 7, 1], "inflow_pressure", True, [(64, 2, 0), (128, 1, 1)]),
    "odd_nx_falls_back": ([23, 6, 5], [23, 9, 1], "noslip", False, [(64, 4, 1)]),
}


@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19", "D3Q27"])
@pytest.mark.parametrize("dt", ["f32", "f16dev", "bf16dev"])
@pytest.mark.parametrize("case", list(PAIR_CASES))
def test_paired_kernel_is_bit_identical_to_one_cell_kernel(dev, lat, dt, case):
    """Two x-adjacent cells per thread (paired loads, cross-lane shifts, extra
    loads at wavefront and row ends) against one cell per thread."""
    import torch
    from tests.gpu_util import bc_tables
    L = O.Lattice(lat)
    n3, n2, bcs, forced, launches = PAIR_CASES[case]
    n = n3 if L.D == 3 else n2
    rng = np.random.default_rng(11)
    f64 = L.w[:, None, None, None] * (1 + 0.05 * rng.standard_normal((L.Q, n[2], n[1], n[0])))
    src = torch.from_numpy(f64).to(dev)
    a = _dev_tensor(torch, L.Q, n, dt, dev)
    fa = abi.field(a, n, 0, L.D, abi.DTYPE[dt])
    abi.convert(abi.field(src, n, 0, L.D), fa, abi.LATTICE[lat])
    g, _, _ = bc_tables(L, bcs)
    F = [2e-4, -1e-4, 5e-5][:L.D] if forced else None
    lid, did = abi.LATTICE[lat], abi.DTYPE[dt]
    outs = []
    try:
        for bx, by, nt in [(64, 4, None)] + launches:
            b = torch.full_like(a, float("nan"))
            if nt is None:
                abi.set_launch_dev(lid, did, bx, by, 1, 0)
            else:
                abi.set_launch_dev(lid, did, bx, by, 2, nt)
            abi.step(fa, abi.field(b, n, 0, L.D, abi.DTYPE[dt]), g, 1.0 / 0.6, F)
            torch.cuda.synchronize()
            outs.append(b.view(torch.int16 if b.element_size() == 2 else torch.int32).cpu().numpy())
    finally:
        abi.set_launch_dev(lid, did)
    for k, o in enumerate(outs[1:]):
        assert np.array_equal(o, outs[0]), (case, launches[k])


@pytest.mark.parametrize("lat", ["D3Q19", "D3Q27"])
@pytest.mark.parametrize("dt", ["f32", "f16dev", "bf16dev"])
def test_row_interleaved_order_gives_the_same_bits(dev, lat, dt):
    """The reduced-storage kernels (paired and one-cell launches) read and
    write the row-interleaved order through strides: the same step from the
    same stored deviations gives identical bits in both orders."""
    import torch
    L = O.Lattice(lat)
    lid, did = abi.LATTICE[lat], abi.DTYPE[dt]
    n = [24, 10, 7]
    rng = np.random.default_rng(17)
    f64 = L.w[:, None, None, None] * (1 + 0.05 * rng.standard_normal((L.Q, n[2], n[1], n[0])))
    src = torch.from_numpy(f64).to(dev)
    a = _dev_tensor(torch, L.Q, n, dt, dev)
    abi.convert(abi.field(src, n, 0, 3), abi.field(a, n, 0, 3, did), lid)
    ar = a.permute(1, 2, 0, 3).contiguous()            # (z, y, q, x)
    g = abi.resolve_bc(lid, [abi.BC["Periodic"]] * 6)
    F = [2e-4, -1e-4, 5e-5]
    try:
        for cfg in ((64, 2, 2, 1), (128, 2, 1, 0)):
            abi.set_launch_dev(lid, did, *cfg)
            b = torch.empty_like(a)
            abi.step(abi.field(a, n, 0, 3, did), abi.field(b, n, 0, 3, did), g, 1.0 / 0.6, F)
            br = torch.empty_like(ar)
            abi.step(abi.field(ar, n, 0, 3, did, layout="row"),
                     abi.field(br, n, 0, 3, did, layout="row"), g, 1.0 / 0.6, F)
            torch.cuda.synchronize()
            assert torch.equal(br.permute(2, 0, 1, 3).contiguous().view(torch.int16 if dt != "f32"
                                                                       else torch.int32),
                               b.view(torch.int16 if dt != "f32" else torch.int32)), cfg
    finally:
        abi.set_launch_dev(lid, did)


STEP2_SIZES = [[13, 7, 6], [40, 33, 70], [1, 5, 3], [64, 16, 64], [130, 9, 5]]


@pytest.mark.parametrize("lat", ["D3Q15", "D3Q19", "D3Q27"])
@pytest.mark.parametrize("dt", ["f32"])
@pytest.mark.parametrize("si", range(len(STEP2_SIZES)))
@pytest.mark.parametrize("forced", [False, True])
def test_step2_equals_two_single_steps(dev, lat, dt, si, forced):
    """FP32-storage mfa_step2 (intermediate state rounded to FP32 in LDS)
    against two mfa_step calls, bit for bit: boxes smaller than its tiles
    (64 x 4, and 32 x 8 for D3Q27), not a multiple of them, and more
    z-planes than one block marches (70 > 64), in both population orders."""
    import torch
    L = O.Lattice(lat)
    lid, did = abi.LATTICE[lat], abi.DTYPE[dt]
    n = STEP2_SIZES[si]
    print(f"case step2_dev lattice={lat} dtype={dt} n={n} forced={forced}", flush=True)
    rng = np.random.default_rng(23 + si)
    f64 = L.w[:, None, None, None] * (1 + 0.05 * rng.standard_normal((L.Q, n[2], n[1], n[0])))
    src = torch.from_numpy(f64).to(dev)
    a = _dev_tensor(torch, L.Q, n, dt, dev)
    abi.convert(abi.field(src, n, 0, 3), abi.field(a, n, 0, 3, did), lid)
    g = abi.resolve_bc(lid, [abi.BC["Periodic"]] * 6)
    F = [2e-4, -1e-4, 5e-5] if forced else None
    omega = 1.0 / 0.6
    word = torch.int16 if dt != "f32" else torch.int32
    mid, two = torch.empty_like(a), torch.empty_like(a)
    abi.step(abi.field(a, n, 0, 3, did), abi.field(mid, n, 0, 3, did), g, omega, F)
    abi.step(abi.field(mid, n, 0, 3, did), abi.field(two, n, 0, 3, did), g, omega, F)
    pair = torch.full_like(a, float("nan"))
    abi.step2(abi.field(a, n, 0, 3, did), abi.field(pair, n, 0, 3, did), g, omega, F)
    ar = a.permute(1, 2, 0, 3).contiguous()            # (z, y, q, x)
    pr = torch.full_like(ar, float("nan"))
    abi.step2(abi.field(ar, n, 0, 3, did, layout="row"), abi.field(pr, n, 0, 3, did, layout="row"),
              g, omega, F)
    torch.cuda.synchronize()
    assert torch.equal(pair.view(word), two.view(word))
    assert torch.equal(pr.permute(2, 0, 1, 3).contiguous().view(word), two.view(word))


def test_step2_reduced_refuses_2d_walls_and_16bit(dev):
    """Where the pass is not offered (2-D, walls, FP16 / BF16 words, which
    measured slower than two single steps) it returns MFA_ERR_UNSUPPORTED."""
    import torch
    from tests.gpu_util import bc_tables
    g27 = abi.resolve_bc(3, [abi.BC["Periodic"]] * 6)
    for dt in ("f16dev", "bf16dev"):
        h = torch.zeros((27, 4, 4, 4), dtype=getattr(torch, TORCH_DT[dt]), device=dev)
        with pytest.raises(abi.MfaError) as e:
            abi.step2(abi.field(h, [4, 4, 4], 0, 3, abi.DTYPE[dt]),
                      abi.field(h.clone(), [4, 4, 4], 0, 3, abi.DTYPE[dt]), g27, 1.0)
        assert e.value.code == -2
    a = torch.zeros((9, 1, 4, 4), dtype=torch.float16, device=dev)
    g = abi.resolve_bc(0, [abi.BC["Periodic"]] * 6)
    with pytest.raises(abi.MfaError) as e:
        abi.step2(abi.field(a, [4, 4], 0, 2, abi.DTYPE["f16dev"]),
                  abi.field(a.clone(), [4, 4], 0, 2, abi.DTYPE["f16dev"]), g, 1.0)
    assert e.value.code == -2
    g3, _, _ = bc_tables(O.Lattice("D3Q27"), "lid")
    b = torch.zeros((27, 4, 4, 4), dtype=torch.float32, device=dev)
    with pytest.raises(abi.MfaError) as e:
        abi.step2(abi.field(b, [4, 4, 4], 0, 3, abi.DTYPE["f32"]),
                  abi.field(b.clone(), [4, 4, 4], 0, 3, abi.DTYPE["f32"]), g3, 1.0)
    assert e.value.code == -2


@pytest.mark.parametrize("bc", ["periodic", "lid", "xper_walls", "inflow_pressure", "inflow_outflow"])
@pytest.mark.parametrize("n", [[13, 7], [130, 9], [1, 5], [70, 33]])
@pytest.mark.parametrize("forced", [False, True])
def test_step2_2d_f32_equals_two_single_steps(dev, bc, n, forced):
    """FP32-storage mfa_step2 on D2Q9 boxes, periodic or with walls (second-
    step wall values from the intermediate deviations in LDS), against two
    mfa_step calls, bit for bit, in both population orders."""
    import torch
    from tests.gpu_util import bc_tables
    L = O.Lattice("D2Q9")
    did = abi.DTYPE["f32"]
    print(f"case step2_2d_f32 bc={bc} n={n} forced={forced}", flush=True)
    rng = np.random.default_rng(31)
    f64 = L.w[:, None, None, None] * (1 + 0.05 * rng.standard_normal((L.Q, 1, n[1], n[0])))
    src = torch.from_numpy(f64).to(dev)
    a = torch.empty((L.Q, 1, n[1], n[0]), dtype=torch.float32, device=dev)
    abi.convert(abi.field(src, n, 0, 2), abi.field(a, n, 0, 2, did), 0)
    g, _, _ = bc_tables(L, bc)
    F = [2e-4, -1e-4] if forced else None
    omega = 1.0 / 0.6
    mid, two = torch.empty_like(a), torch.empty_like(a)
    abi.step(abi.field(a, n, 0, 2, did), abi.field(mid, n, 0, 2, did), g, omega, F)
    abi.step(abi.field(mid, n, 0, 2, did), abi.field(two, n, 0, 2, did), g, omega, F)
    pair = torch.full_like(a, float("nan"))
    abi.step2(abi.field(a, n, 0, 2, did), abi.field(pair, n, 0, 2, did), g, omega, F)
    ar = a.permute(1, 2, 0, 3).contiguous()
    pr = torch.full_like(ar, float("nan"))
    abi.step2(abi.field(ar, n, 0, 2, did, layout="row"), abi.field(pr, n, 0, 2, did, layout="row"), g, omega, F)
    torch.cuda.synchronize()
    assert torch.equal(pair.view(torch.int32), two.view(torch.int32))
    assert torch.equal(pr.permute(2, 0, 1, 3).contiguous().view(torch.int32), two.view(torch.int32))
