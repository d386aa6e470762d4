"""Batched engine: every FP64 ensemble member is identical to the same
lattice advanced alone by mfa_step (itself identical to the oracle); masked
members match the oracle's obstacle step; graph replay equals eager
launches; retirement freezes a lattice; residuals match numpy."""
import numpy as np
import pytest

from muphfasa_amd import abi, ensemble as E
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _single(dev, nx, ny, omega, f0, steps, bc):
    import torch
    L = O.Lattice("D2Q9")
    a = torch.from_numpy(np.ascontiguousarray(f0)).to(dev)
    b = torch.empty_like(a)
    fa, fb = abi.field(a, [nx, ny], 0, 2), abi.field(b, [nx, ny], 0, 2)
    for _ in range(steps):
        abi.step(fa, fb, bc, omega)
        a, fa, b, fb = b, fb, a, fa
    torch.cuda.synchronize()
    return a.cpu().numpy()


@pytest.mark.parametrize("tile", [(64, 4), (16, 16), (256, 1)])
@pytest.mark.parametrize("graph", [False, True])
def test_members_identical_to_single_lattice(dev, tile, graph):
    import torch
    sizes = [(33, 20), (64, 64), (17, 40), (100, 7)]
    omegas = [1.0 / 0.6, 1.0 / 0.9, 1.0 / 0.53, 1.0]
    ens = E.Ensemble(sizes, omegas, tile=tile, grid=3 if graph else 0)
    ens.init()
    f0 = [ens.lattice_populations(l).copy() for l in range(len(sizes))]
    ens.run(60, graph=graph)
    torch.cuda.synchronize()
    bc = E.lid_cavity_bc()
    for l, (nx, ny) in enumerate(sizes):
        ref = _single(dev, nx, ny, omegas[l], f0[l], 60, bc)
        assert np.array_equal(ens.lattice_populations(l), ref), l


def test_member_matches_oracle_and_moments(dev):
    import torch
    L = O.Lattice("D2Q9")
    sizes = [(24, 18), (40, 40)]
    omegas = [1.0 / 0.7, 1.0 / 0.55]
    ens = E.Ensemble(sizes, omegas)
    ens.init()
    ens.run(200)
    ens.moments()
    torch.cuda.synchronize()
    fu = np.zeros((6, 3))
    fu[3, 0] = 0.1
    res = O.resolve_bc(2, [1, 1, 1, 0], np.ones(6), fu)
    for l, (nx, ny) in enumerate(sizes):
        f0, _, _ = O.init_equilibrium(L, np.ones((1, ny, nx)), np.zeros((2, 1, ny, nx)), [nx, ny])
        O.apply_bc(L, f0, [nx, ny], [0, 0, 0], res)
        fe = O.run(L, f0, [nx, ny], omegas[l], None, [0, 0], res, 200)
        assert np.array_equal(ens.lattice_populations(l), O.valid(L, fe))
        rho, vel = O.moments(L, fe, [nx, ny])
        r, v = ens.lattice_fields(l)
        assert np.array_equal(r, rho[0]) and np.array_equal(v, vel[:, 0])


@pytest.mark.parametrize("tile", [(64, 4), (16, 16), (32, 2)])
def test_masked_member_matches_oracle(dev, tile):
    """Obstacles inside tiles, on tile edges and only in a tile's one-cell
    rim (the near flag must catch pulls across tile edges); a lattice with no
    solid at all runs the unmasked path."""
    import torch
    L = O.Lattice("D2Q9")
    sizes = [(30, 26), (20, 20), (40, 12)]
    masks = [np.zeros((26, 30), np.uint8), np.zeros((20, 20), np.uint8),
             np.zeros((12, 40), np.uint8)]
    masks[0][8:14, 10:19] = 1
    masks[0][20, 16] = 1          # a lone solid cell on a tile row edge
    masks[1][3:5, 15:17] = 1
    omegas = [1.0 / 0.6, 1.0 / 0.8, 1.0 / 0.7]
    ens = E.Ensemble(sizes, omegas, masks=masks, tile=tile)
    ens.init()
    ens.run(150)
    torch.cuda.synchronize()
    fu = np.zeros((6, 3))
    fu[3, 0] = 0.1
    res = O.resolve_bc(2, [1, 1, 1, 0], np.ones(6), fu)
    for l, (nx, ny) in enumerate(sizes):
        f0 = O.new_fab(L, [nx, ny])
        O.valid(L, f0)[...] = L.w[:, None, None, None]
        O.apply_bc(L, f0, [nx, ny], [0, 0, 0], res)
        fe = O.run_masked(L, f0, [nx, ny], omegas[l], [0, 0], res, masks[l][None], 150)
        assert np.array_equal(ens.lattice_populations(l), O.valid(L, fe)), l


def test_retirement_and_residual(dev):
    import torch
    sizes = [(32, 32), (48, 24), (16, 64)]
    ens = E.Ensemble(sizes, [1.0 / 0.6] * 3)
    ens.init()
    ens.run(100)
    ens.moments(ens.vel_prev)
    ens.run(100)
    r = ens.residual()
    v_now = ens.vel.cpu().numpy()
    torch.cuda.synchronize()
    # recompute lattice 1's residual with numpy from two moment snapshots
    assert np.all(np.isfinite(r)) and np.all(r > 0)
    frozen = ens.lattice_populations(1).copy()
    ens.set_active([1, 0, 1])
    ens.run(50)
    torch.cuda.synchronize()
    assert np.array_equal(ens.lattice_populations(1), frozen)
    # residual of an unchanged field is zero
    ens.set_active([1, 1, 1])
    ens.moments(ens.vel_prev)
    r2 = ens.residual()
    assert np.all(r2 == 0.0)
    assert v_now.shape == (2 * ens.total_cells,)


def test_fp32_member_close_to_fp64(dev):
    import torch
    sizes = [(64, 64)]
    out = {}
    for dt in ("f64", "f32"):
        ens = E.Ensemble(sizes, [1.0 / 0.56], dtype=dt)
        ens.init()
        ens.run(2000)
        ens.moments()
        torch.cuda.synchronize()
        out[dt] = ens.lattice_fields(0)[1]
    err = np.abs(out["f32"] - out["f64"]).max() / 0.1
    assert err < 1e-5, err


def test_warm_init_average_matches_oracle(dev):
    """mfa_ensemble_init_warm: feq of the given fields (rest state in solid
    cells), one masked step, then the mean of the two states, against the
    oracle's equilibrium initialisation and obstacle step."""
    import torch
    L = O.Lattice("D2Q9")
    sizes = [(30, 26), (20, 22)]
    masks = [np.zeros((26, 30), np.uint8), np.zeros((22, 20), np.uint8)]
    masks[0][8:14, 10:19] = 1
    masks[1][3:6, 12:16] = 1
    omegas = [1.0 / 0.6, 1.0 / 0.53]
    rng = np.random.default_rng(8)
    rhos, vels = [], []
    for (nx, ny), m in zip(sizes, masks):
        r = 1.0 + 0.01 * rng.standard_normal((ny, nx))
        v = 0.05 * rng.standard_normal((2, ny, nx))
        r[m == 1] = 1.0
        v[:, m == 1] = 0.0
        rhos.append(r)
        vels.append(v)
    ens = E.Ensemble(sizes, omegas, masks=masks)
    rho_d = torch.from_numpy(np.concatenate([r.ravel() for r in rhos])).to(dev)
    vel_d = torch.from_numpy(np.concatenate([np.concatenate([v[0].ravel() for v in vels]),
                                             np.concatenate([v[1].ravel() for v in vels])])).to(dev)
    ens.init(rho_d, vel_d, average=True)
    torch.cuda.synchronize()
    fu = np.zeros((6, 3))
    fu[3, 0] = 0.1
    res = O.resolve_bc(2, [1, 1, 1, 0], np.ones(6), fu)
    for l, (nx, ny) in enumerate(sizes):
        f0, _, _ = O.init_equilibrium(L, rhos[l][None], vels[l][:, None], [nx, ny])
        O.apply_bc(L, f0, [nx, ny], [0, 0, 0], res)
        f1 = O.run_masked(L, f0, [nx, ny], omegas[l], [0, 0], res, masks[l][None], 1)
        expect = 0.5 * (O.valid(L, f0) + O.valid(L, f1))
        assert np.array_equal(ens.lattice_populations(l), expect), l


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_neq_start_matches_the_chapman_enskog_statement(dev, dtype):
    """mfa_ensemble_init_start(MFA_START_NEQ): feq of the given fields plus
    -3 (tau - 1) w rho Q:S, against tests/ce_ref.py (the same statement in
    numpy) on two masked cavities with the lid moving; mass and momentum
    of every fluid cell are those of (rho, u) (the added part has neither).
    With MFA_START_AVERAGE it is followed by one step and the mean of the
    two states, as the plain warm start."""
    import torch
    import ce_ref
    sizes = [(30, 26), (20, 22)]
    masks = [np.zeros((26, 30), np.uint8), np.zeros((22, 20), np.uint8)]
    masks[0][8:14, 10:19] = 1
    masks[1][3:6, 12:16] = 1
    masks[1][10, 0:3] = 1
    omegas = [1.0 / 0.6, 1.0 / 0.53]
    rng = np.random.default_rng(9)
    rhos, vels = [], []
    for (nx, ny), m in zip(sizes, masks):
        yy, xx = np.meshgrid(np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
        r = 1.0 + 0.01 * np.sin(xx / 3.0) * np.cos(yy / 5.0)
        v = 0.05 * np.stack([np.sin(yy / 4.0) + 0.3 * np.cos(xx / 2.0), np.cos(xx / 3.0) * np.sin(yy / 7.0)])
        v += 1e-3 * rng.standard_normal(v.shape)
        r[m == 1] = 1.0
        v[:, m == 1] = 0.0
        rhos.append(r)
        vels.append(v)
    ens = E.Ensemble(sizes, omegas, dtype=dtype, masks=masks)
    rho_d = torch.from_numpy(np.concatenate([r.ravel() for r in rhos])).to(dev)
    vel_d = torch.from_numpy(np.concatenate([np.concatenate([v[0].ravel() for v in vels]),
                                             np.concatenate([v[1].ravel() for v in vels])])).to(dev)
    ens.init(rho_d, vel_d, neq=True)
    torch.cuda.synchronize()
    walls = [(0.0, 0.0), (0.0, 0.0), (0.0, 0.0), (0.1, 0.0)]
    tol = 1e-15 if dtype == "f64" else 2e-7
    for l in range(2):
        want = ce_ref.ce_start(rhos[l], vels[l], masks[l], omegas[l], walls)
        got = ens.lattice_populations(l)[:, 0]
        if dtype == "f32":
            got = got + ce_ref.W[:, None, None]
        assert np.abs(got - want).max() < tol, (l, np.abs(got - want).max())
        fl = masks[l] == 0
        neq = want - ce_ref.feq(rhos[l], vels[l])
        assert np.abs(neq[:, fl]).max() > 1e-4                # the part is there ...
        assert np.abs(neq.sum(axis=0)[fl]).max() < 1e-15      # ... and carries no mass
        assert np.abs(np.einsum("qa,qyx->ayx", ce_ref.C, neq)[:, fl]).max() < 1e-15
    if dtype == "f64":
        f0 = [ens.lattice_populations(l).copy() for l in range(2)]
        ens2 = E.Ensemble(sizes, omegas, dtype=dtype, masks=masks)
        ens2.init(rho_d, vel_d, neq=True, average=True)
        ens.step()
        torch.cuda.synchronize()
        for l in range(2):
            f1 = ens.lattice_populations(l)          # the state after one step
            assert np.array_equal(ens2.lattice_populations(l), 0.5 * (f0[l] + f1)), l


def _masked_moments_reference(L, fab, n, mask):
    """rho, u of the pulled populations with half-way bounce-back at solid
    neighbours (a population pulled from a solid cell is the cell's own
    opposite one), wall ghosts from the filled fab; sums in population order
    like the reference's computeDensity / computeVelocity."""
    nx, ny = n
    v = fab[:, 0]                                  # (Q, ny+2, nx+2) with ghosts
    own = v[:, 1:-1, 1:-1]
    rho = np.zeros((ny, nx))
    jx = np.zeros((ny, nx))
    jy = np.zeros((ny, nx))
    pulled = []
    for q in range(L.Q):
        cx, cy = int(L.c[q, 0]), int(L.c[q, 1])
        f = v[q, 1 - cy:1 - cy + ny, 1 - cx:1 - cx + nx].copy()
        sx = np.arange(nx)[None, :] - cx
        sy = np.arange(ny)[:, None] - cy
        inside = (sx >= 0) & (sx < nx) & (sy >= 0) & (sy < ny)
        solid_src = np.zeros((ny, nx), bool)
        solid_src[inside] = mask[np.clip(sy, 0, ny - 1), np.clip(sx, 0, nx - 1)][inside]
        f[solid_src] = own[L.opp[q]][solid_src]
        pulled.append(f)
    for q in range(L.Q):
        rho = rho + pulled[q]
    for q in range(L.Q):
        if L.c[q, 0] > 0:
            jx = jx + pulled[q]
        elif L.c[q, 0] < 0:
            jx = jx - pulled[q]
        if L.c[q, 1] > 0:
            jy = jy + pulled[q]
        elif L.c[q, 1] < 0:
            jy = jy - pulled[q]
    u = np.stack([jx / rho, jy / rho])
    rho[mask == 1] = 1.0
    u[:, mask == 1] = 0.0
    return rho, u


@pytest.mark.parametrize("tile", [(64, 4), (16, 16)])
def test_masked_moments_match_reference(dev, tile):
    """The moments behind the residual and the converged fields of obstacle
    cavities: identical to a numpy restatement of the pull with obstacle
    bounce-back, on tiles next to solids and on tiles away from them."""
    import torch
    L = O.Lattice("D2Q9")
    sizes = [(30, 26), (40, 12)]
    masks = [np.zeros((26, 30), np.uint8), np.zeros((12, 40), np.uint8)]
    masks[0][8:14, 10:19] = 1
    masks[0][20, 16] = 1
    omegas = [1.0 / 0.6, 1.0 / 0.7]
    ens = E.Ensemble(sizes, omegas, masks=masks, tile=tile)
    ens.init()
    ens.run(120)
    ens.moments()
    torch.cuda.synchronize()
    fu = np.zeros((6, 3))
    fu[3, 0] = 0.1
    res = O.resolve_bc(2, [1, 1, 1, 0], np.ones(6), fu)
    for l, (nx, ny) in enumerate(sizes):
        fab = O.new_fab(L, [nx, ny])
        O.valid(L, fab)[...] = ens.lattice_populations(l)
        O.apply_bc(L, fab, [nx, ny], [0, 0, 0], res)
        rho_e, u_e = _masked_moments_reference(L, fab, [nx, ny], masks[l])
        rho_g, u_g = ens.lattice_fields(l)
        assert np.array_equal(rho_g, rho_e), l
        assert np.array_equal(u_g, u_e), l


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("masked", [False, True])
def test_step2_equals_two_steps(dev, dt, masked):
    """mfa_ensemble_step2 (two steps per launch, intermediate state in LDS,
    second-step wall values and obstacle bounce-back from it) against two
    mfa_ensemble_step launches, bit for bit: lattices narrower than the
    32 x 16 pair tile, not a multiple of it, one cell wide, with obstacles at
    random cells (also on tile edges)."""
    import torch
    sizes = [(40, 23), (64, 64), (1, 5), (97, 33), (33, 1), (70, 17)]
    omegas = [1 / 0.56, 1 / 0.7, 1 / 0.9, 1 / 0.6, 1 / 0.8, 1 / 0.65]
    rng = np.random.default_rng(29)
    masks = None
    if masked:
        masks = [(rng.random((ny, nx)) < 0.1).astype(np.uint8) for nx, ny in sizes]
    ens = [E.Ensemble(sizes, omegas, dtype=dt, masks=masks) for _ in range(2)]
    total = ens[0].total_cells
    rho = torch.from_numpy(1.0 + 0.01 * rng.standard_normal(total)).to(dev)
    vel = torch.from_numpy(0.05 * rng.standard_normal(2 * total)).to(dev)
    for e in ens:
        e.init(rho, vel)
    for _ in range(5):
        ens[0].step(pair=True)
    for _ in range(10):
        ens[1].step()
    torch.cuda.synchronize()
    word = torch.int64 if dt == "f64" else torch.int32
    assert torch.equal(ens[0].fa.view(word), ens[1].fa.view(word))
    for e in ens:
        e.close()
