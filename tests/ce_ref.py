"""numpy statement of the first-order Chapman-Enskog start of the batched
engine (mfa_ensemble_init_start with MFA_START_NEQ), for D2Q9 lattices whose
four faces are half-way walls (the lid cavities of the ensemble) and whose
solid cells are walls at rest:

  f_q = feq_q(rho, u) - 3 (tau - 1) w_q rho Q_q:S,   Q_q = c_q c_q - I/3,

with S the symmetric part of grad u by central differences, second-order
one-sided differences through a half-way wall (the quadratic through the wall
velocity at 1/2 and the two cells on the other side), and the difference
across the gap when both neighbours are walls.  Test infrastructure only.
"""
import numpy as np

W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
C = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1]], float)


def feq(rho, u):
    """(9, ny, nx) equilibrium in the reference's order of operations."""
    usq = u[0] * u[0] + u[1] * u[1]
    out = np.empty((9,) + rho.shape)
    for q in range(9):
        cu = C[q, 0] * u[0] + C[q, 1] * u[1] if q else 0 * u[0]
        out[q] = W[q] * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq)
    return out


def strain(u, mask, wall_u, periodic=(False, False)):
    """S (2, 2, ny, nx).  wall_u[face] = (ux, uy) for faces 0..3 = x-lo,
    x-hi, y-lo, y-hi; solid cells (mask 1) are walls at rest; a periodic
    axis wraps instead of ending in walls."""
    ny, nx = u.shape[1:]
    g = np.zeros((2, 2, ny, nx))
    for y in range(ny):
        for x in range(nx):
            if mask[y, x]:
                continue
            for a in range(2):
                kinds, vals = [], []
                for s in (-1, 1):
                    xx, yy = (x + s, y) if a == 0 else (x, y + s)
                    if periodic[a]:
                        xx, yy = xx % nx, yy % ny
                    inside = 0 <= xx < nx and 0 <= yy < ny
                    if not inside:
                        face = 2 * a + (1 if s > 0 else 0)
                        kinds.append("w")
                        vals.append(np.asarray(wall_u[face], float))
                    elif mask[yy, xx]:
                        kinds.append("w")
                        vals.append(np.zeros(2))
                    else:
                        kinds.append("f")
                        vals.append(u[:, yy, xx])
                m, p, c = vals[0], vals[1], u[:, y, x]
                if kinds == ["f", "f"]:
                    d = 0.5 * (p - m)
                elif kinds == ["f", "w"]:
                    d = 4 / 3 * p - c - m / 3
                elif kinds == ["w", "f"]:
                    d = p / 3 + c - 4 / 3 * m
                else:
                    d = p - m
                g[a, :, y, x] = d
    return 0.5 * (g + g.transpose(1, 0, 2, 3))


def ce_start(rho, u, mask, omega, wall_u, periodic=(False, False)):
    """(9, ny, nx) populations; solid cells hold the rest equilibrium."""
    S = strain(u, mask, wall_u, periodic)
    tau = 1.0 / omega
    f = feq(rho, u)
    tr3 = (S[0, 0] + S[1, 1]) / 3
    for q in range(9):
        qs = C[q, 0] ** 2 * S[0, 0] + 2 * C[q, 0] * C[q, 1] * S[0, 1] + C[q, 1] ** 2 * S[1, 1] - tr3
        f[q] = f[q] - 3 * (tau - 1) * W[q] * rho * qs
    solid = mask.astype(bool)
    f[:, solid] = W[:, None]
    return f
