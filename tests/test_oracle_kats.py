"""Pins the CPU oracle against the reference's own known-answer tests.

Each block restates one reference unit/regression test (cited) and checks the
oracle against the closed forms those tests use.  The checks are chosen so a
plausible slip in the oracle (wrong sign, wrong opposite index, dropped half
force, pull vs push) fails at least one of them; test_oracle_mutations.py
shows that by planting such slips in a copy of the oracle's formulas.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O

LATS = ["D2Q9", "D3Q15", "D3Q19", "D3Q27"]


# ---- tests/unit/test_lattice.cpp:22-137 ------------------------------------
@pytest.mark.parametrize("name", LATS)
def test_lattice_invariants(name):
    L = O.Lattice(name)
    assert abs(L.w.sum() - 1.0) < 1e-12                                   # :22-30
    for q in range(L.Q):
        o = L.opp[q]
        assert L.opp[o] == q and L.w[o] == L.w[q]                        # :33-46
        assert np.all(L.c[o] == -L.c[q])                                  # :49-61
    assert len({tuple(c) for c in L.c}) == L.Q                            # :64-76
    D = L.D
    m1 = (L.w[:, None] * L.c[:, :D]).sum(0)
    assert np.all(np.abs(m1) < 1e-12)                                     # :79-90
    m2 = np.einsum("q,qa,qb->ab", L.w, L.c[:, :D].astype(float), L.c[:, :D].astype(float))
    assert np.allclose(m2, np.eye(D) / 3.0, atol=1e-12)                   # :91-104
    if D == 2:
        assert np.all(L.c[:, 2] == 0)


# ---- tests/unit/test_equilibrium.cpp:66-466 ---------------------------------
def _vel_samples(D):
    if D == 2:
        return [(0.0, 0.0), (0.05, 0.0), (0.0, -0.07), (0.05, -0.02), (0.10, 0.05),
                (-0.08, 0.06), (0.02, 0.02)]
    return [(0.0, 0.0, 0.0), (0.05, 0.0, 0.0), (0.0, -0.06, 0.0), (0.0, 0.0, 0.04),
            (0.04, -0.02, 0.03), (0.08, 0.05, -0.03), (-0.05, 0.04, 0.06)]


@pytest.mark.parametrize("name", LATS)
def test_equilibrium_moments(name):
    L = O.Lattice(name)
    rng = np.random.default_rng(12345)
    samples = [(r, u) for r in (0.5, 1.0, 1.3, 2.0) for u in _vel_samples(L.D)]
    samples += [(rng.uniform(0.8, 1.2), tuple(rng.uniform(-0.06, 0.06, L.D))) for _ in range(20)]
    for rho, u in samples:
        f = np.array([O.feq(L, rho, u, q) for q in range(L.Q)])
        tol = 1e-12 * max(1.0, rho)
        assert abs(f.sum() - rho) < tol                                  # mass
        j = (f[:, None] * L.c[:, :L.D]).sum(0)
        assert np.allclose(j, rho * np.array(u), atol=tol)               # momentum
        if not any(u):
            assert np.allclose(f, L.w * rho, atol=1e-15)                 # rest state
        if max(abs(x) for x in u) < 0.15:
            assert np.all(f > 0)                                         # positivity
        r2, u2 = O.cell_moments(L, f)                                    # round trip
        assert abs(r2 - rho) < tol and np.allclose(u2, u, atol=1e-12)
        # second moment of the 2nd-order equilibrium: rho(u u + cs2 I)
        c = L.c[:, :L.D].astype(float)
        P = np.einsum("q,qa,qb->ab", f, c, c)
        assert np.allclose(P, rho * (np.outer(u, u) + np.eye(L.D) / 3.0), atol=1e-12)


# ---- tests/unit/test_forcing.cpp:112-233 -------------------------------------
@pytest.mark.parametrize("name", LATS)
def test_guo_forcing_identities(name):
    L = O.Lattice(name)
    D = L.D
    rng = np.random.default_rng(12345)
    c = L.c[:, :D].astype(float)
    for omega in (0.8, 1.0, 1.5):
        u = rng.uniform(-0.05, 0.05, D)
        F = rng.uniform(-0.05, 0.05, D) * 1e-2
        Fi = np.array([O.guo(L, q, u, F, omega) for q in range(L.Q)])
        pref = 1.0 - 0.5 * omega
        assert abs(Fi.sum()) < 1e-14                                     # :133-150
        assert np.allclose((Fi[:, None] * c).sum(0), pref * F, atol=1e-14)
        m2 = np.einsum("q,qa,qb->ab", Fi, c, c)
        assert np.allclose(m2, pref * (np.outer(u, F) + np.outer(F, u)), atol=1e-14)
    for omega in (0.8, 1.0, 1.5):                                        # :161-199
        rho = 1.0 + 0.1 * rng.uniform(-0.05, 0.05)
        u0 = rng.uniform(-0.05, 0.05, D)
        F = rng.uniform(-0.05, 0.05, D) * 1e-2
        f = np.array([O.feq(L, rho, u0, q) for q in range(L.Q)])
        rc, uc = O.cell_moments(L, f, F)
        assert abs(rc - rho) < 1e-14
        assert np.allclose(uc, u0 + F / (2 * rho), atol=1e-14)
        fe = np.array([O.feq(L, rho, uc, q) for q in range(L.Q)])
        fo = np.array([O.bgk(f[q], fe[q], omega) + O.guo(L, q, uc, F, omega)
                       for q in range(L.Q)])
        assert abs(fo.sum() - rho) < 1e-14
        assert np.allclose((fo[:, None] * c).sum(0), rho * u0 + F, atol=1e-14)


@pytest.mark.parametrize("name", LATS)
def test_forced_step_with_zero_force_is_bit_identical(name):
    """tests/unit/test_forcing.cpp:201-233 at field level: F = 0 through the
    forced path equals the plain path bit for bit."""
    L = O.Lattice(name)
    n = [5, 4, 3] if L.D == 3 else [6, 5]
    rng = np.random.default_rng(7)
    fin = O.new_fab(L, n, 0.0)
    fin[...] = L.w[:, None, None, None] * (1 + 0.05 * rng.standard_normal(fin.shape))
    a = O.new_fab(L, n, 0.0)
    b = O.new_fab(L, n, 0.0)
    O.step(L, fin, a, n, 1.3)
    O.step(L, fin, b, n, 1.3, F=[0.0] * L.D)
    assert np.array_equal(O.valid(L, a), O.valid(L, b))


# ---- pull direction (src/LBM/Streaming.H:20-35) ------------------------------
@pytest.mark.parametrize("name", LATS)
def test_stream_is_pull(name):
    """A unit pulse of population q at cell x arrives at x + c_q."""
    L = O.Lattice(name)
    n = [5, 5, 5] if L.D == 3 else [5, 5]
    x0 = np.array([2, 2, 2 if L.D == 3 else 0])
    for q in range(L.Q):
        f = O.new_fab(L, n, 0.0)
        O.valid(L, f)[q, x0[2], x0[1], x0[0]] = 1.0
        g = O.new_fab(L, n, 0.0)
        O.stream(L, f, g, n)
        v = O.valid(L, g)[q]
        dst = x0 + L.c[q]
        assert v[dst[2], dst[1], dst[0]] == 1.0 and v.sum() == 1.0


# ---- tests/regression/uniform_force/main.cpp:47-77 (checked every step) ------
@pytest.mark.parametrize("name,n", [("D2Q9", [8, 8]), ("D3Q19", [6, 6, 6]), ("D3Q27", [4, 5, 6])])
def test_uniform_force_closed_form(name, n):
    L = O.Lattice(name)
    D = L.D
    F = np.array([4e-4, -2e-4, 1e-4])[:D]
    nn = O.n3(n)
    rho = np.ones(nn[::-1])
    vel = np.zeros((D,) + tuple(nn[::-1]))
    f, rho1, vel1 = O.init_equilibrium(L, rho, vel, n, F)
    assert np.allclose(vel1, 0.0, atol=1e-15)                            # reported u0 = 0
    per = [1, 1, 1][:D]
    res = O.resolve_bc(D, [O.BC["Periodic"]] * (2 * D))
    O.fill_periodic(L, f, n, per + [0] * (3 - D))
    for k in range(1, 31):
        f = O.run(L, f, n, 1.0 / 0.8, F, per, res, 1)
        r, u = O.moments(L, f, n, F)
        for d in range(D):
            assert np.max(np.abs(u[d] - k * F[d] / 1.0)) < 1e-12
            assert np.ptp(u[d]) < 1e-12
        assert np.max(np.abs(r - 1.0)) < 1e-12


# ---- tests/unit/test_bc_corners.cpp ------------------------------------------
def _pattern_fab(L, n):
    f = O.new_fab(L, n, np.nan)
    v = O.valid(L, f)
    nn = O.n3(n)
    k, j, i = np.meshgrid(range(nn[2]), range(nn[1]), range(nn[0]), indexing="ij")
    for q in range(L.Q):
        if L.D == 2:
            v[q] = i * 1000 + j * 10 + q * 0.1 + 0.5                     # :173
        else:
            v[q] = i * 1000 + j * 10 + k * 0.1 + q * 0.01 + 0.5          # :184
    return f


def _noslip(D):
    return O.resolve_bc(D, [O.BC["NoSlip"]] * (2 * D))


def _fab_get(L, f, i, j, k, q):
    if L.D == 2:
        return f[q, 0, j + 1, i + 1]
    return f[q, k + 1, j + 1, i + 1]


def _ghosts(L, n):
    nn = O.n3(n)
    rng = [range(-1, nn[d] + 1) if d < L.D else range(0, 1) for d in range(3)]
    for k in rng[2]:
        for j in rng[1]:
            for i in rng[0]:
                if 0 <= i < nn[0] and 0 <= j < nn[1] and 0 <= k < nn[2]:
                    continue
                yield i, j, k


def _inside(n, p):
    nn = O.n3(n)
    return all(0 <= p[d] < nn[d] for d in range(3))


@pytest.mark.parametrize("name", LATS)
def test_bc_source_cell_noslip(name):
    """test_bc_corners.cpp:476-548 / :617-668: ghost f(g,n) == f(g+c_n, opp n)."""
    L = O.Lattice(name)
    n = [6, 6] if L.D == 2 else [4, 4, 4]
    f = _pattern_fab(L, n)
    O.apply_bc(L, f, n, [0, 0, 0], _noslip(L.D))
    checked = 0
    for g in _ghosts(L, n):
        for q in range(L.Q):
            src = np.array(g) + L.c[q]
            if not _inside(n, src):
                continue
            exp = _fab_get(L, f, *src, L.opp[q])
            got = _fab_get(L, f, *g, q)
            assert abs(got - exp) <= 1e-12
            checked += 1
    assert checked > 0
    # corner needs: 2-D (-1,-1) -> exactly (1,1); D3Q19 (-1,-1,-1) -> none
    corner = np.array([-1, -1, -1 if L.D == 3 else 0])
    need = [q for q in range(L.Q) if _inside(n, corner + L.c[q])]
    if name == "D2Q9":
        assert len(need) == 1 and tuple(L.c[need[0]][:2]) == (1, 1)
    if name == "D3Q19":
        assert need == []
    if name == "D3Q27":
        assert len(need) == 1


def test_bc_moving_wall_correction_2d():
    """test_bc_corners.cpp:550-609: lid ghost (2,6) gets -2 rho w_opp (c_opp.U)/cs2."""
    L = O.Lattice("D2Q9")
    n = [6, 6]
    f = _pattern_fab(L, n)
    ft = [O.BC["NoSlip"]] * 4
    ft[3] = O.BC["MovingWall"]
    fu = np.zeros((6, 3))
    fu[3, 0] = 0.1
    res = O.resolve_bc(2, ft, np.ones(6), fu)
    O.apply_bc(L, f, n, [0, 0, 0], res)
    g = np.array([2, 6, 0])
    for q in range(L.Q):
        src = g + L.c[q]
        if not _inside(n, src):
            continue
        m = L.opp[q]
        f_src = src[0] * 1000 + src[1] * 10 + m * 0.1 + 0.5
        corr = -2.0 * 1.0 * L.w[m] * (L.c[m, 0] * 0.1) * 3.0
        assert abs(_fab_get(L, f, 2, 6, 0, q) - (f_src + corr)) <= 1e-12
    # Appendix D: all four corners resolve to NoSlip (rank 3 beats 2)
    rtype = res[0]
    for idx in (0, 2, 6, 8):
        assert rtype[idx] == O.BC["NoSlip"]
    assert rtype[1 + 3 * 2] == O.BC["MovingWall"]


@pytest.mark.parametrize("name", LATS)
@pytest.mark.parametrize("periodic_x", [0, 1])
def test_bc_coverage(name, periodic_x):
    """test_bc_corners.cpp:352-433: after periodic fill + BC every pull source
    of every valid cell is finite."""
    L = O.Lattice(name)
    n = [4, 4] if L.D == 2 else [4, 4, 4]
    rng = np.random.default_rng(42)
    f = O.new_fab(L, n, np.nan)
    O.valid(L, f)[...] = rng.uniform(0.5, 1.5, O.valid(L, f).shape)
    per = [periodic_x, 0, 0]
    ft = [O.BC["Periodic"] if (periodic_x and i < 2) else O.BC["NoSlip"] for i in range(2 * L.D)]
    O.fill_periodic(L, f, n, per)
    O.apply_bc(L, f, n, per, O.resolve_bc(L.D, ft))
    g = O.new_fab(L, n, np.nan)
    O.stream(L, f, g, n)
    assert np.all(np.isfinite(O.valid(L, g)))


def test_bc_priority_table():
    """test_bc_corners.cpp:712-823."""
    B = O.BC
    assert O.has_conflict([B["MovingWall"], B["MovingWall"]])
    assert O.has_conflict([B["Inflow"], B["Outflow"]])
    assert O.has_conflict([B["Inflow"], B["PressureOutlet"]])
    assert O.has_conflict([B["Outflow"], B["PressureOutlet"]])
    assert not O.has_conflict([B["NoSlip"], B["PressureOutlet"]])
    assert O.has_conflict([B["MovingWall"]] * 3)
    # NoSlip beats MovingWall at the xlo/yhi corner; velocity is NoSlip's (0)
    fu = np.zeros((6, 3))
    fu[3, 0] = 0.1
    rt, rr, ru = O.resolve_bc(2, [B["NoSlip"], B["NoSlip"], B["NoSlip"], B["MovingWall"]],
                              np.ones(6), fu)
    assert rt[0 + 3 * 2] == B["NoSlip"] and ru[0 + 3 * 2, 0] == 0.0
    # MovingWall beats PressureOutlet; U is the wall's
    rt, rr, ru = O.resolve_bc(2, [B["NoSlip"], B["PressureOutlet"], B["MovingWall"], B["NoSlip"]],
                              np.ones(6), np.array([[0, 0, 0]] * 2 + [[0.1, 0, 0]] + [[0, 0, 0]] * 3))
    assert rt[2 + 3 * 0] == B["MovingWall"] and ru[2, 0] == 0.1
    # conflicting faces abort (resolve raises)
    with pytest.raises(ValueError):
        O.resolve_bc(2, [B["Inflow"], B["Outflow"], B["NoSlip"], B["Inflow"]])
    # periodic dims drop out of a code: x-periodic, y walls -> corner code uses ylo
    rt, _, _ = O.resolve_bc(2, [B["Periodic"], B["Periodic"], B["NoSlip"], B["MovingWall"]])
    assert rt[0] == B["NoSlip"] and rt[2 + 3 * 2] == B["MovingWall"]
    assert rt[0 + 3 * 1] == B["Periodic"]


@pytest.mark.parametrize("name", LATS)
def test_pressure_outlet_kernel(name):
    """test_bc_corners.cpp:834-902: anti-bounce-back pins the face density."""
    L = O.Lattice(name)
    D = L.D
    n = [4, 4, 4][:D]
    vel = np.array([0.03, -0.02, 0.01])[:D]
    ft = [O.BC["NoSlip"]] * (2 * D)
    ft[1] = O.BC["PressureOutlet"]
    res = O.resolve_bc(D, ft, np.ones(6), None)
    g = (4, 1, 1 if D == 3 else 0)
    for rho in (1.0, 1.15):
        f = O.new_fab(L, n, np.nan)
        for q in range(L.Q):
            O.valid(L, f)[q] = O.feq(L, rho, vel, q)
        O.apply_bc(L, f, n, [0, 0, 0], res)
        for q in range(L.Q):
            src = np.array(g) + L.c[q]
            if not _inside(n, src):
                continue
            got = _fab_get(L, f, *g, q)
            if rho == 1.0:
                assert abs(got - O.feq(L, 1.0, vel, q)) < 1e-14
            else:
                m = L.opp[q]
                pair = got + _fab_get(L, f, *src, m)
                ref = O.feq(L, 1.0, vel, q) + O.feq(L, 1.0, vel, m)
                assert abs(pair - ref) < 1e-13


# ---- tests/unit/test_mass_conservation.cpp:130-199 (fixed: moments refreshed) --
@pytest.mark.parametrize("name,n", [("D2Q9", [8, 8]), ("D2Q9", [12, 6]), ("D3Q15", [6, 6, 6]),
                                    ("D3Q19", [6, 6, 6]), ("D3Q27", [6, 6, 6])])
def test_closed_box_mass_conservation(name, n):
    L = O.Lattice(name)
    D = L.D
    nn = O.n3(n)
    rng = np.random.default_rng(3)
    rho = 1.0 + 0.01 * rng.standard_normal(nn[::-1])
    vel = 0.02 * rng.standard_normal((D,) + tuple(nn[::-1]))
    f, _, _ = O.init_equilibrium(L, rho, vel, n)
    res = _noslip(D)
    O.apply_bc(L, f, n, [0, 0, 0], res)
    r0, _ = O.moments(L, f, n, None, streamed=True)
    f2 = O.run(L, f, n, 1.0, None, [0, 0, 0][:D], res, 50)
    r1, v1 = O.moments(L, f2, n, None, streamed=True)
    assert abs(r1.sum() - r0.sum()) / r0.sum() < 1e-12
    assert np.abs(v1).max() > 0                      # the flow actually evolved


# ---- tests/regression/lid2d/main.cpp:45-54,297-340 ----------------------------
def _extremum(fv, sign):
    n = len(fv)
    m = int(np.argmax(sign * fv))
    val = fv[m]
    if 0 < m < n - 1:
        den = fv[m - 1] - 2.0 * fv[m] + fv[m + 1]
        if den != 0.0:
            delta = 0.5 * (fv[m - 1] - fv[m + 1]) / den
            val = fv[m] - 0.25 * (fv[m - 1] - fv[m + 1]) * delta
    return val


@pytest.mark.slow
def test_lid2d_marchi_extrema():
    """N=64 D2Q9 Re=100 tau=0.692 cavity, 20001 steps: centreline extrema
    within 1 % of Marchi et al. (2009), and the measured table in
    tests/regression/lid2d/README.md:68-72 to 5e-5."""
    L = O.Lattice("D2Q9")
    N = 64
    n = [N, N]
    u_lid = 0.1
    tau = 0.692
    ft = [O.BC["NoSlip"], O.BC["NoSlip"], O.BC["NoSlip"], O.BC["MovingWall"]]
    fu = np.zeros((6, 3))
    fu[3, 0] = u_lid
    res = O.resolve_bc(2, ft, np.ones(6), fu)
    f, _, _ = O.init_equilibrium(L, np.ones((1, N, N)), np.zeros((2, 1, N, N)), n)
    O.apply_bc(L, f, n, [0, 0, 0], res)
    f = O.run(L, f, n, 1.0 / tau, None, [0, 0], res, 20001)
    rho, vel = O.moments(L, f, n)
    u = vel[0, 0] / u_lid
    v = vel[1, 0] / u_lid
    c0, c1 = N // 2 - 1, N // 2
    u_of_y = 0.5 * u[:, c0] + 0.5 * u[:, c1]
    v_of_x = 0.5 * v[c0, :] + 0.5 * v[c1, :]
    umin, vmax, vmin = _extremum(u_of_y, -1), _extremum(v_of_x, 1), _extremum(v_of_x, -1)
    for got, ref in ((umin, -2.140417e-1), (vmax, 1.79572814e-1), (vmin, -2.53804e-1)):
        assert abs(got - ref) / abs(ref) < 0.01
    for got, ref in ((umin, -0.21393), (vmax, 0.17919), (vmin, -0.25262)):
        assert abs(got - ref) < 5e-5
    assert abs(rho.sum() - N * N) / (N * N) < 1e-9


# ---- solid obstacles (engine extension; no reference counterpart) --------------
def test_masked_step_reduces_to_plain_step_without_solids():
    L = O.Lattice("D2Q9")
    n = [9, 7]
    rng = np.random.default_rng(4)
    f = O.new_fab(L, n, 0.0)
    f[...] = L.w[:, None, None, None] * (1 + 0.05 * rng.standard_normal(f.shape))
    a, b = O.new_fab(L, n, 0.0), O.new_fab(L, n, 0.0)
    O.step(L, f, a, n, 1.3)
    O.step_masked(L, f, b, n, 1.3, np.zeros((1, 7, 9), np.uint8))
    assert np.array_equal(O.valid(L, a), O.valid(L, b))


def test_masked_step_bounces_back_and_conserves_mass():
    L = O.Lattice("D2Q9")
    n = [12, 10]
    mask = np.zeros((1, 10, 12), np.uint8)
    mask[0, 3:6, 4:7] = 1
    rng = np.random.default_rng(6)
    f = O.new_fab(L, n, np.nan)
    O.valid(L, f)[...] = L.w[:, None, None, None] * (1 + 0.05 * rng.standard_normal((9, 1, 10, 12)))
    res = O.resolve_bc(2, [O.BC["NoSlip"]] * 4)
    O.apply_bc(L, f, n, [0, 0, 0], res)
    out = O.new_fab(L, n, np.nan)
    O.step_masked(L, f, out, n, 1.0, mask)
    # omega = 1: out = feq of the pulled set.  For the fluid cell (3,4) left of
    # the block, populations q=3,6,7 (c_x = -1) come from solid cells and are
    # the cell's own opposite populations 1,8,5.
    fl = [f[q, 0, 4 + 1 - L.c[q, 1], 3 + 1 - L.c[q, 0]] for q in range(9)]
    for q in (3, 6, 7):
        fl[q] = f[L.opp[q], 0, 5, 4]
    rho, u = O.cell_moments(L, np.array(fl))
    expect = [O.feq(L, rho, u, q) for q in range(9)]
    assert np.allclose(out[:, 0, 5, 4], expect, rtol=0, atol=1e-15)
    # mass over fluid cells is conserved by a closed box with obstacles
    fm = f.copy()
    for q in range(9):
        O.valid(L, fm)[q][mask.astype(bool)] = L.w[q]
    fluid = ~mask.astype(bool)
    m0 = O.moments(L, fm, n, streamed=False)[0][fluid].sum()
    fe = O.run_masked(L, fm, n, 1.0 / 0.7, [0, 0], res, mask, 40)
    r1 = O.moments(L, fe, n, streamed=False)[0]
    assert abs(r1[fluid].sum() - m0) / m0 < 1e-12


@pytest.mark.parametrize("lat,n,mgs", [("D3Q19", [20, 18, 22], 8), ("D2Q9", [37, 21], 16),
                                       ("D3Q27", [9, 10, 11], 4), ("D3Q15", [7, 5, 6], 64)])
def test_boxed_run_equals_single_box(lat, n, mgs):
    """The CPU baseline's box layout (boxes of at most max_grid_size cells,
    each with its own ghost layer, filled from the neighbouring boxes) gives
    the single box's bits after the same steps: ragged last boxes, a box
    bigger than the domain, forced."""
    L = O.Lattice(lat)
    rng = np.random.default_rng(3)
    f = O.new_fab(L, n)
    v = O.valid(L, f)
    v[...] = L.w[:, None, None, None] * (1 + 0.05 * rng.standard_normal(v.shape))
    F = [1e-5, -2e-6, 3e-6] if L.D == 3 else [1e-5, -2e-6]
    per = [1] * L.D + [0] * (3 - L.D)
    ref = O.resolve_bc(L.D, [O.BC["Periodic"]] * (2 * L.D))
    g = f.copy()
    O.fill_periodic(L, g, n, per)
    expect = O.run(L, g, n, 1.0 / 0.8, F, per, ref, 5)
    b = f.copy()
    assert O.time_run_boxes(L, b, n, mgs, 1.0 / 0.8, F, 2, 3) >= 0
    assert np.array_equal(O.valid(L, b), O.valid(L, expect))
