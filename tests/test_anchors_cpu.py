"""The oracle against the closed forms the reference's regression suite checks
(plane Couette start-up and grid convergence, 2-D Taylor-Green decay and grid
convergence), with the reference's CI thresholds:
.github/workflows/couette-regression.yml:73,77,142,146 and
.github/workflows/tgv2d-regression.yml:81.  These pin the oracle's
collision, streaming and wall treatment to textbook solutions, independently
of any parity with the GPU path."""
import os

import anchors as A

INP = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inputs")


def _read(name):
    return open(os.path.join(INP, name)).read()


def test_oracle_couette_startup_2d():
    err, rate_err = A.couette_startup(A.oracle_snapshots, _read("couette_2d"))
    print(err, rate_err)
    assert err <= 0.01 and rate_err <= 0.01


def test_oracle_couette_startup_3d_d3q19():
    err, rate_err = A.couette_startup(A.oracle_snapshots, _read("couette_3d"))
    print(err, rate_err)
    assert err <= 0.01 and rate_err <= 0.01


def test_oracle_couette_convergence_2d():
    errs = A.couette_convergence(A.oracle_snapshots, _read("couette_2d"), [16, 32, 64])
    orders = A.observed_orders(errs)
    print(errs, orders)
    assert min(orders) >= 1.8


def test_oracle_tgv2d_convergence_and_decay():
    text = _read("tgv2d")
    errs = A.tgv2d_convergence(A.oracle_snapshots, text, [32, 64, 128])
    orders = A.observed_orders(errs)
    err, rate_err = A.tgv2d_decay(A.oracle_snapshots, text)
    print(errs, orders, err, rate_err)
    assert min(orders) >= 1.8
    assert err <= 0.01 and rate_err <= 0.01


def test_a_wall_left_at_rest_fails_the_couette_check():
    """The measure is not vacuous: with the upper wall a NoSlip face (its
    bc.hiy_Uwall then ignored, as buildBCManager does) the fluid stays at rest
    and the profile error against the moving-wall series is of order one."""
    text = _read("couette_2d").replace("bc.hi_type = Periodic  MovingWall",
                                       "bc.hi_type = Periodic  NoSlip")
    assert A.parse_inputs(text)["bc.hi_type"] == ["Periodic", "NoSlip"]
    err, rate_err = A.couette_startup(A.oracle_snapshots, text, every=500)
    print(err, rate_err)
    assert err > 0.5
