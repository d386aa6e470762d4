"""The oracle's known-answer checks have teeth: a copy of the cell formulas
(taken from the oracle, in Python) with one plausible slip planted at a time
must fail the same identities that tests/test_oracle_kats.py applies to the
oracle, while the unmodified copy passes and agrees with the oracle."""
import numpy as np
import pytest

from oracle import oracle as O

L = O.Lattice("D3Q19")
C = L.c[:, :3].astype(float)


def feq(rho, u, q, slip=None):
    cu = C[q] @ u
    usq = u @ u
    a, b, c = 3.0, 4.5, 1.5
    if slip == "feq_sign":
        a = -3.0
    if slip == "feq_half":
        b = 3.0
    if slip == "feq_usq":
        c = 0.5
    return L.w[q] * rho * (1 + a * cu + b * cu * cu - c * usq)


def guo(q, u, F, omega, slip=None):
    cF, cu, uF = C[q] @ F, C[q] @ u, u @ F
    pref = 1.0 - 0.5 * omega
    if slip == "guo_pref":
        pref = 1.0
    if slip == "guo_drop_cucf":
        return pref * L.w[q] * 3.0 * (cF - uF)
    return pref * L.w[q] * 3.0 * (cF + 3.0 * cu * cF - uF)


def contrib(m, rho, U, slip=None):
    cu = C[m] @ U
    if slip == "bb_sign":
        return 2.0 * rho * L.w[m] * cu * 3.0
    return -2.0 * rho * L.w[m] * cu * 3.0


def opp(q, slip=None):
    if slip == "opp_wrong":
        return L.opp[q] if q % 2 == 0 else q     # odd populations not reversed
    return L.opp[q]


def identities_hold(slip):
    rng = np.random.default_rng(0)
    ok = True
    for _ in range(5):
        rho = 1 + 0.1 * rng.uniform(-1, 1)
        u = rng.uniform(-0.05, 0.05, 3)
        F = rng.uniform(-1e-3, 1e-3, 3)
        omega = 1.3
        f = np.array([feq(rho, u, q, slip) for q in range(L.Q)])
        ok &= abs(f.sum() - rho) < 1e-12
        ok &= np.allclose(f @ C, rho * u, atol=1e-12)
        ok &= np.allclose(np.einsum("q,qa,qb->ab", f, C, C), rho * (np.outer(u, u) + np.eye(3) / 3),
                          atol=1e-12)
        g = np.array([guo(q, u, F, omega, slip) for q in range(L.Q)])
        ok &= abs(g.sum()) < 1e-14
        ok &= np.allclose(g @ C, (1 - 0.5 * omega) * F, atol=1e-14)
        ok &= np.allclose(np.einsum("q,qa,qb->ab", g, C, C),
                          (1 - 0.5 * omega) * (np.outer(u, F) + np.outer(F, u)), atol=1e-14)
        # moving wall: the bounced population carries momentum toward the wall
        # velocity; bounce-back of a rest equilibrium at a wall moving with U
        # must give the equilibrium at U to first order: f(g,q) - w_q = 3 w_q c_q.U
        U = np.array([0.1, 0.0, 0.0])
        for q in range(1, L.Q):
            m = opp(q, slip)
            if not np.array_equal(C[m], -C[q]):
                ok = False
                break
            val = L.w[m] + contrib(m, 1.0, U, slip)
            ok &= abs((val - L.w[q]) - 3 * L.w[q] * (C[q] @ U) * 2) < 1e-15 or \
                abs((val - L.w[q]) - 6 * L.w[q] * (C[q] @ U)) < 1e-15
    return bool(ok)


def test_unmodified_copy_passes_and_matches_oracle():
    assert identities_hold(None)
    rng = np.random.default_rng(3)
    u = rng.uniform(-0.05, 0.05, 3)
    F = rng.uniform(-1e-3, 1e-3, 3)
    for q in range(L.Q):
        assert abs(feq(1.1, u, q) - O.feq(L, 1.1, u, q)) < 1e-16
        assert abs(guo(q, u, F, 1.3) - O.guo(L, q, u, F, 1.3)) < 1e-18
        assert abs(contrib(q, 1.0, [0.1, 0, 0]) - O.bb_contrib(L, q, 1.0, [0.1, 0, 0])) < 1e-18


@pytest.mark.parametrize("slip", ["feq_sign", "feq_half", "feq_usq", "guo_pref", "guo_drop_cucf",
                                  "bb_sign", "opp_wrong"])
def test_each_planted_slip_is_caught(slip):
    assert not identities_hold(slip)
