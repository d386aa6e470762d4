"""float32 restatement, in numpy, of the reduced-storage kernels' collision
(deviation form d_q = f_q - w_q, BGK + Guo): the same operations in the same
order, so a kernel result differs from it only by float32 evaluation noise
and the final rounding to the storage type.  Test infrastructure."""
import numpy as np


def _fma32(a, b, c):
    """float32 fused multiply-add: one rounding of the exact a*b + c (the
    float64 product of two float32 values is exact; the float64 sum is
    rounded twice, which differs from one rounding only in rare ties)."""
    return np.asarray(np.float64(a) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(
        np.float32)


def _q_of(L, cx, cy, cz):
    for q in range(L.Q):
        c = L.c[q]
        if c[0] == cx and c[1] == cy and (c[2] if L.D == 3 else 0) == cz:
            return q
    return -1


def _moments_f32(L, p):
    """drho and j summed over velocity columns: c_z first, then c_y, then
    c_x (the kernels' order)."""
    def col(cx, cy):
        qp, q0, qm = _q_of(L, cx, cy, 1), _q_of(L, cx, cy, 0), _q_of(L, cx, cy, -1)
        if qp >= 0 and q0 >= 0:
            return (p[qm] + p[q0]) + p[qp]
        if qp >= 0:
            return p[qm] + p[qp]
        return p[q0]
    C, Dy, j2 = [], [], None
    for cx in (-1, 0, 1):
        am, a0, ap = col(cx, -1), col(cx, 0), col(cx, 1)
        C.append((am + a0) + ap)
        Dy.append(ap - am)
        for cy in (-1, 0, 1):
            qp, qm = _q_of(L, cx, cy, 1), _q_of(L, cx, cy, -1)
            if L.D == 3 and qp >= 0:
                b = p[qp] - p[qm]
                j2 = b if j2 is None else j2 + b
    drho = (C[0] + C[1]) + C[2]
    j0 = C[2] - C[0]
    j1 = (Dy[0] + Dy[1]) + Dy[2]
    return drho, [j0, j1, j2 if j2 is not None else np.zeros_like(j0)]


def _cdot32(L, q, a):
    r = None
    for k in range(L.D):
        c = L.c[q, k]
        if c != 0:
            t = a[k] if c > 0 else -a[k]
            r = t if r is None else r + t
    return r


def ref_step_f32(L, d, omega, F, periodic=True):
    """float32 deviation-form pull + BGK (+Guo) on a periodic box, in the
    kernels' arithmetic: column-factored moments, and per pair q, opp(q)
    the shared even part E = drho + rho (4.5 (c.u)^2 - 1.5 u.u) and odd part
    O = 3 rho c.u, with fused multiply-adds."""
    f32 = np.float32
    Q, D = L.Q, L.D
    pulled = np.empty_like(d)
    for q in range(Q):
        c = L.c[q]
        pulled[q] = np.roll(d[q], shift=(c[2], c[1], c[0]), axis=(0, 1, 2))
    drho, j = _moments_f32(L, pulled)
    rho = f32(1) + drho
    inv = f32(1) / rho
    Fv = [f32(F[a]) if (F is not None and a < len(F)) else f32(0) for a in range(3)]
    u = [(_fma32(f32(0.5), Fv[a], j[a]) if F is not None else j[a]) * inv for a in range(D)]
    if D == 2:
        u.append(np.zeros_like(u[0]))
    usq = _fma32(u[1], u[1], u[0] * u[0])
    uF = _fma32(u[1], Fv[1], u[0] * Fv[0])
    if D == 3:
        usq = _fma32(u[2], u[2], usq)
        uF = _fma32(u[2], Fv[2], uF)
    om = f32(omega)
    omw = f32(1) - om
    r3 = f32(3) * rho
    m15usq = f32(-1.5) * usq
    pref = f32(1) - f32(0.5) * om
    out = np.empty_like(d)
    for q in range(Q):
        m = L.opp[q]
        w = f32(L.w[q])
        if m == q:
            fe = w * _fma32(rho, m15usq, drho)
            o = _fma32(om, fe, omw * pulled[q])
            if F is not None:
                o = _fma32(pref * w * f32(3), -uF, o)
            out[q] = o
        elif q < m:
            cu = _cdot32(L, q, u)
            E = _fma32(rho, _fma32(f32(4.5) * cu, cu, m15usq), drho)
            O = r3 * cu
            oq = _fma32(om, w * (E + O), omw * pulled[q])
            omm = _fma32(om, w * (E - O), omw * pulled[m])
            if F is not None:
                cF = _cdot32(L, q, Fv)
                kg = pref * w * f32(3)
                ge = _fma32(f32(3) * cu, cF, -uF)
                oq = _fma32(kg, ge + cF, oq)
                omm = _fma32(kg, ge - cF, omm)
            out[q] = oq
            out[m] = omm
    return out
