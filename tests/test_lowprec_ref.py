"""The float32 restatement of the reduced-storage collision
(tests/lowprec_ref.py), which the GPU tests hold the kernels to, against the
FP64 oracle step on the same populations.  The restatement factors the
update (column sums for the moments; the even and odd parts of each pair q,
opp(q); fused multiply-adds), so this pins the factoring: it must agree with
the reference's term-by-term update to float32 rounding, while a wrong sign
or a dropped term shows as an error of order w |c.u| or w |c.F|."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.lowprec_ref import ref_step_f32


@pytest.mark.parametrize("lat", ["D2Q9", "D3Q15", "D3Q19", "D3Q27"])
@pytest.mark.parametrize("forced", [False, True])
def test_float32_restatement_matches_fp64_oracle(lat, forced):
    L = O.Lattice(lat)
    n = [12, 7, 5] if L.D == 3 else [12, 7, 1]
    rng = np.random.default_rng(5)
    w = L.w[:, None, None, None]
    d = (w * 0.05 * rng.standard_normal((L.Q, n[2], n[1], n[0]))).astype(np.float32)
    F = [2e-4, -1e-4, 5e-5][:L.D] if forced else None
    omega = 1.0 / 0.6
    got = ref_step_f32(L, d, omega, F)
    f = O.new_fab(L, n)
    O.valid(L, f)[...] = w + d.astype(np.float64)
    O.fill_periodic(L, f, n, [1] * L.D + [0] * (3 - L.D))
    out = O.new_fab(L, n)
    O.step(L, f, out, n, omega, F)
    expect = O.valid(L, out) - w
    err = (np.abs(got - expect) / w).max()
    assert err < 1e-6, err
    # the terms the factoring must keep are far above that floor
    assert (np.abs(expect - d) / w).max() > 1e-3
