"""The batched and precision workloads at their own scale.

* 512 D2Q9 lid cavities (side 64..256, Re log-uniform in [50, 1000], lid
  u = 0.1, seed 2024: the batched workload), 4 000 steps in one ensemble:
  every cavity keeps its mass (FP64), sampled members are bit-identical to
  the same lattice run alone through mfa_step, and the FP32-deviation
  ensemble stays within a stated distance of the FP64 one.
* D3Q27 Taylor-Green vortex at tau = 0.51146 (the low-viscosity regime of
  the 384^3, Re = 800 study) on a 128^3 box for 5 000 steps, stored as FP64,
  FP32, FP16 and BF16 deviations: every run stays finite and its velocity
  and kinetic energy stay within the stated distance of the FP64 run.
The FP64 arithmetic these are measured against is the reference's forced
and plain BGK step (src/LBM/Collision.cpp:18-93)."""
import math

import numpy as np
import pytest

from muphfasa_amd import abi, ensemble as E, tgv

pytestmark = pytest.mark.gpu

STEPS3 = 4000


def _single(dev, nx, ny, omega, f0, steps, bc):
    import torch
    a = torch.from_numpy(np.ascontiguousarray(f0)).to(dev)
    b = torch.empty_like(a)
    fa, fb = abi.field(a, [nx, ny], 0, 2), abi.field(b, [nx, ny], 0, 2)
    for _ in range(steps):
        abi.step(fa, fb, bc, omega)
        a, fa, b, fb = b, fb, a, fa
    torch.cuda.synchronize()
    return a.cpu().numpy()


def _masses(ens):
    cells = [a * b for a, b in ens.sizes]
    return np.array([float(s.sum()) for s in ens.rho.split(cells)])


@pytest.fixture(scope="module")
def config3():
    N, Re, tau = E.cavity_ensemble_params(512)
    return [(int(x), int(x)) for x in N], 1.0 / tau


def test_512_cavities_fp64_mass_and_members(dev, config3):
    import torch
    sizes, omegas = config3
    ens = E.Ensemble(sizes, omegas, dtype="f64")
    ens.init()
    ens.moments()
    m0 = _masses(ens)
    pick = sorted({int(np.argmin([s[0] for s in sizes])), int(np.argmax([s[0] for s in sizes])),
                   int(np.argmin(omegas)), int(np.argmax(omegas)), 17, 300})
    f0 = {l: ens.lattice_populations(l).copy() for l in pick}
    ens.run(STEPS3, graph=True)
    ens.moments()
    torch.cuda.synchronize()
    m1 = _masses(ens)
    drift = np.abs(m1 - m0) / m0
    print(f"512 cavities, {STEPS3} steps: max mass drift {drift.max():.2e}", flush=True)
    assert drift.max() < 1e-12
    assert torch.isfinite(ens.vel).all().item()
    bc = E.lid_cavity_bc()
    for l in pick:
        nx, ny = sizes[l]
        ref = _single(dev, nx, ny, omegas[l], f0[l], STEPS3, bc)
        assert np.array_equal(ens.lattice_populations(l), ref), (l, sizes[l], omegas[l])


def test_512_cavities_fp32_against_fp64(dev, config3):
    """The FP32-deviation ensemble after 4 000 steps (the start-up transient):
    max over every cell of every cavity of |u_f32 - u_f64| / U below 1e-4
    (measured 2.6e-5, median over the cavities 1.2e-5)."""
    import torch
    sizes, omegas = config3
    vel = {}
    for dt in ("f64", "f32"):
        ens = E.Ensemble(sizes, omegas, dtype=dt)
        ens.init()
        ens.run(STEPS3, graph=True)
        ens.moments()
        torch.cuda.synchronize()
        vel[dt] = ens.vel.clone()
        del ens
        torch.cuda.empty_cache()
    d = (vel["f32"] - vel["f64"]).abs() / 0.1
    cells = [a * b for a, b in sizes]
    tc = sum(cells)
    per = torch.maximum(d[:tc], d[tc:]).split(cells)
    per = np.array([float(p.max()) for p in per])
    print(f"FP32 vs FP64 after {STEPS3} steps: max {per.max():.2e}, median {np.median(per):.2e}",
          flush=True)
    assert per.max() < 1e-4


# bounds on the FP64-relative L2 error of u and of the kinetic energy after
# 5 000 steps at tau = 0.51146 on 128^3, per storage type (measured: FP32
# 1.7e-4 / 1.1e-4, FP16 7.9e-4 / 1.5e-4, BF16 3.2e-2 / 1.7e-2)
TGV_BOUNDS = {"f32": (1e-3, 1e-3), "f16dev": (5e-3, 5e-3), "bf16dev": (0.1, 0.1)}


def test_tgv_d3q27_low_viscosity_all_storages(dev):
    N, u0, tau = 128, 0.05, 0.51146
    Re = 3.0 * u0 * (N / (2 * math.pi)) / (tau - 0.5)
    r = tgv.run(N=N, u0=u0, Re=Re, steps=5000, sample=1000)
    assert abs(r["tau"] - tau) < 1e-12
    res = r["results"]
    assert res["f64"]["mass_rel_drift"] < 1e-12
    ke = [e for _, e in res["f64"]["ke"]]
    assert all(b < a for a, b in zip(ke, ke[1:]))          # the vortex decays
    for st, (bu, bk) in TGV_BOUNDS.items():
        out = res[st]
        assert out["finite"], st
        l2 = out["l2_u_err_vs_f64"][-1][1]
        print(f"{st}: L2(u) error {l2:.2e}, KE error {out['ke_rel_err_final']:.2e}, "
              f"mass drift {out['mass_rel_drift']:.2e}, {out['mlups']} MLUPS", flush=True)
        assert l2 < bu and out["ke_rel_err_final"] < bk, st
