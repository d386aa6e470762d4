"""Several boxes of one lattice on one device (include/mfa_level.h): the ghost
exchange between boxes (FillBoundary with periodic images), the physical
boundary fill of each box at its place in the domain, and the step of every
box through its ghosts (mfa_pull_collide per box, and mfa_level_step for all
boxes in one launch, on alternate steps), against the oracle on the whole
domain as one box, bit for bit.  The box layouts are uneven (as AMReX's
max_grid_size chopping leaves them) and include boxes of one cell."""
import numpy as np
import pytest

from muphfasa_amd import abi
from oracle import oracle as O
from tests.gpu_util import bc_tables, random_state, to_dev, to_host

pytestmark = pytest.mark.gpu

# cuts along each dimension: box edges
CUTS = {
    2: ([0, 12, 30], [0, 8, 15, 21], [0, 1]),
    3: ([0, 9, 22], [0, 6, 11, 17], [0, 1, 7, 13]),
}


def _boxes(D):
    cx, cy, cz = CUTS[D]
    lo, n = [], []
    for k in range(len(cz) - 1):
        for j in range(len(cy) - 1):
            for i in range(len(cx) - 1):
                lo.append([cx[i], cy[j], cz[k] if D == 3 else 0])
                n.append([cx[i + 1] - cx[i], cy[j + 1] - cy[j], (cz[k + 1] - cz[k]) if D == 3 else 1])
    dom = [cx[-1], cy[-1], cz[-1] if D == 3 else 1]
    return dom, lo, n


def _box_fab(L, full_valid, lo, n):
    """Oracle-layout fab (ngrow 1, NaN ghosts) holding the box's valid cells."""
    fab = O.new_fab(L, n[:L.D])
    v = O.valid(L, fab)
    if L.D == 3:
        v[...] = full_valid[:, lo[2]:lo[2] + n[2], lo[1]:lo[1] + n[1], lo[0]:lo[0] + n[0]]
    else:
        v[...] = full_valid[:, :, lo[1]:lo[1] + n[1], lo[0]:lo[0] + n[0]]
    return fab


def _gather(L, dom, lo, n, fabs):
    out = np.full((L.Q, dom[2], dom[1], dom[0]), np.nan)
    for b, fab in enumerate(fabs):
        v = O.valid(L, fab)
        if L.D == 3:
            out[:, lo[b][2]:lo[b][2] + n[b][2], lo[b][1]:lo[b][1] + n[b][1], lo[b][0]:lo[b][0] + n[b][0]] = v
        else:
            out[:, :, lo[b][1]:lo[b][1] + n[b][1], lo[b][0]:lo[b][0] + n[b][0]] = v
    return out


@pytest.mark.parametrize("lat,bc", [("D2Q9", "lid"), ("D2Q9", "inflow_pressure"), ("D2Q9", "periodic"),
                                    ("D3Q19", "periodic"), ("D3Q19", "lid"), ("D3Q19", "xper_walls"),
                                    ("D3Q27", "inflow_outflow"), ("D3Q15", "noslip")])
@pytest.mark.parametrize("forced", [False, True])
def test_level_steps_equal_one_box_oracle(dev, lat, bc, forced):
    import torch
    L = O.Lattice(lat)
    dom, lo, n = _boxes(L.D)
    print(f"case level lattice={lat} bc={bc} forced={forced} boxes={len(lo)}", flush=True)
    g, ref, per = bc_tables(L, bc)
    F = ([2e-5, -1e-5, 3e-6][:L.D]) if forced else None
    omega = 1.0 / 0.7
    steps = 3
    f0 = random_state(L, dom[:L.D], seed=7)
    # oracle: the whole domain as one box
    f = f0.copy()
    for _ in range(steps):
        O.fill_periodic(L, f, dom[:L.D], per)
        O.apply_bc(L, f, dom[:L.D], per, ref)
        out = O.new_fab(L, dom[:L.D])
        O.step(L, f, out, dom[:L.D], omega, F)
        f = out
    expect = O.valid(L, f)
    # the level
    full0 = O.valid(L, f0)
    ta = [to_dev(_box_fab(L, full0, lo[b], n[b]), dev) for b in range(len(lo))]
    tb = [torch.full_like(t, float("nan")) for t in ta]
    fa = [abi.field(t, n[b][:L.D], 1, L.D) for b, t in enumerate(ta)]
    fb = [abi.field(t, n[b][:L.D], 1, L.D) for b, t in enumerate(tb)]
    lv = abi.Level(abi.LATTICE[lat], dom, per, lo, n)
    try:
        lv.bind(0, fa)
        lv.bind(1, fb)
        cur, nxt = (0, fa), (1, fb)
        for s in range(steps):
            lv.fill_boundary(cur[0])
            lv.apply_bc(cur[0], g)
            if s % 2:        # one launch for all boxes
                lv.step(cur[0], nxt[0], omega, F)
            else:            # mfa_pull_collide box by box
                for b in range(len(lo)):
                    abi.pull_collide(cur[1][b], nxt[1][b], abi.LATTICE[lat], omega, F)
            cur, nxt = nxt, cur
        # moments through the ghosts of the final state
        lv.fill_boundary(cur[0])
        lv.apply_bc(cur[0], g)
        rho_t, vel_t = [], []
        for b in range(len(lo)):
            nb = n[b]
            r = torch.zeros((1, nb[2], nb[1], nb[0]), dtype=torch.float64, device=dev)
            v = torch.zeros((L.D, nb[2], nb[1], nb[0]), dtype=torch.float64, device=dev)
            abi.moments(cur[1][b], abi.field(r, nb[:L.D], 0, L.D), abi.field(v, nb[:L.D], 0, L.D), g, F,
                        streamed="halo")
            rho_t.append(r)
            vel_t.append(v)
        torch.cuda.synchronize()
    finally:
        lv.close()
    final = ta if cur[0] == 0 else tb
    got = _gather(L, dom, lo, n, [to_host(t) for t in final])
    assert np.array_equal(got, expect)
    # moments: oracle's streamed moments of the final state (ghosts filled on the one box)
    O.fill_periodic(L, f, dom[:L.D], per)
    O.apply_bc(L, f, dom[:L.D], per, ref)
    rho_e, vel_e = O.moments(L, f, dom[:L.D], F, streamed=True)
    rho_e = np.asarray(rho_e).reshape(dom[2], dom[1], dom[0])
    vel_e = np.asarray(vel_e).reshape(L.D, dom[2], dom[1], dom[0])
    for b in range(len(lo)):
        l, m = lo[b], n[b]
        sl = (slice(l[2], l[2] + m[2]), slice(l[1], l[1] + m[1]), slice(l[0], l[0] + m[0]))
        assert np.array_equal(to_host(rho_t[b])[0], rho_e[sl]), b
        assert np.array_equal(to_host(vel_t[b]), vel_e[(slice(None),) + sl]), b


@pytest.mark.parametrize("lat,bc", [("D3Q19", "xper_walls"), ("D3Q27", "periodic"), ("D2Q9", "inflow_outflow")])
def test_fill_boundary_writes_exactly_the_periodic_images(dev, lat, bc):
    """After fill_boundary alone, every ghost cell of every box that lies in
    the periodic-grown domain holds its image's valid value (all
    components); every other ghost cell is untouched (NaN)."""
    import torch
    L = O.Lattice(lat)
    dom, lo, n = _boxes(L.D)
    g, _, per = bc_tables(L, bc)
    full = O.valid(L, random_state(L, dom[:L.D], seed=3))
    ta = [to_dev(_box_fab(L, full, lo[b], n[b]), dev) for b in range(len(lo))]
    lv = abi.Level(abi.LATTICE[lat], dom, per, lo, n)
    try:
        lv.bind(0, [abi.field(t, n[b][:L.D], 1, L.D) for b, t in enumerate(ta)])
        lv.fill_boundary(0)
        torch.cuda.synchronize()
        written = lv.ghost_cells
    finally:
        lv.close()
    count = 0
    for b in range(len(lo)):
        fab = to_host(ta[b])
        e = [n[b][d] + 2 if d < L.D else 1 for d in range(3)]
        for z in range(e[2]):
            for y in range(e[1]):
                for x in range(e[0]):
                    p = [x - 1, y - 1, z - 1 if L.D == 3 else 0]
                    local_valid = all(0 <= p[d] < n[b][d] for d in range(3))
                    if local_valid:
                        continue
                    P = [lo[b][d] + p[d] for d in range(3)]
                    inside = all(per[d] or 0 <= P[d] < dom[d] for d in range(L.D))
                    cell = fab[:, z, y, x]
                    if inside:
                        S = [P[d] % dom[d] for d in range(3)]
                        assert np.array_equal(cell, full[:, S[2], S[1], S[0]]), (b, p)
                        count += 1
                    else:
                        assert np.all(np.isnan(cell)), (b, p)
    assert count == written


def test_level_rejects_bad_layouts_and_bindings(dev):
    import torch
    with pytest.raises(abi.MfaError):
        abi.Level(abi.LATTICE["D2Q9"], [8, 8], [1, 1], [[0, 0], [4, 0]], [[5, 8], [4, 8]])  # overlap
    with pytest.raises(abi.MfaError):
        abi.Level(abi.LATTICE["D2Q9"], [8, 8], [1, 1], [[0, 0]], [[8, 7]])                  # gap
    lv = abi.Level(abi.LATTICE["D2Q9"], [8, 8], [1, 1], [[0, 0], [4, 0]], [[4, 8], [4, 8]])
    try:
        t = torch.zeros((9, 1, 10, 7), dtype=torch.float64, device=dev)     # 5 x 8 box: wrong extents
        with pytest.raises(abi.MfaError):
            lv.bind(0, [abi.field(t, [5, 8], 1, 2), abi.field(t, [5, 8], 1, 2)])
        with pytest.raises(abi.MfaError):
            lv.fill_boundary(1)                                                  # slot never bound
        # the boxes of one slot, and an input and an output box, must not share memory
        u = torch.zeros((2, 9, 1, 10, 6), dtype=torch.float64, device=dev)
        shared = abi.field(u[0], [4, 8], 1, 2)
        with pytest.raises(abi.MfaError, match="overlap"):
            lv.bind(0, [shared, shared])
        flat = torch.zeros(2 * 9 * 10 * 6 + 1, dtype=torch.float64, device=dev)
        half = flat[:9 * 10 * 6].view(9, 1, 10, 6)
        skew = flat[9 * 10 * 6 - 6:2 * 9 * 10 * 6 - 6].view(9, 1, 10, 6)    # one row into `half`
        with pytest.raises(abi.MfaError, match="overlap"):
            lv.bind(0, [abi.field(half, [4, 8], 1, 2), abi.field(skew, [4, 8], 1, 2)])
        lv.bind(0, [abi.field(u[0], [4, 8], 1, 2), abi.field(u[1], [4, 8], 1, 2)])
        w = torch.zeros((9, 1, 10, 6), dtype=torch.float64, device=dev)
        lv.bind(1, [abi.field(w, [4, 8], 1, 2), abi.field(u[1], [4, 8], 1, 2)])  # box 1 in both slots
        with pytest.raises(abi.MfaError, match="overlap"):
            lv.step(0, 1, 1.0)
    finally:
        lv.close()


@pytest.mark.parametrize("layout", ["component", "row"])
def test_level_run_bit_identical_in_both_population_orders(dev, layout):
    """muphfasa_amd.level.run: a 40^3 periodic forced domain as 27 uneven
    boxes (max_grid 16: pieces of 14, 13, 13), stepped through graph-replayed
    level steps, equals the one-box mfa_step run bit for bit, with the boxes
    in the MultiFab's component-major order or in the row-interleaved order."""
    from muphfasa_amd import level
    print(f"case level.run layout={layout}", flush=True)
    out = level.run(n=40, max_grid=16, steps=6, dev=dev, layout=layout)
    assert out["level"]["boxes"] == 27 and out["bit_identical"]
    assert out["level_direct"]["bit_identical"]


@pytest.mark.parametrize("lat", ["D2Q9", "D3Q15", "D3Q19", "D3Q27"])
@pytest.mark.parametrize("forced", [False, True])
@pytest.mark.parametrize("launch", [(2, 0), (1, 2), (4, 1)])
def test_direct_level_steps_equal_one_box_oracle(dev, lat, forced, launch):
    """mfa_level_step_direct on the uneven periodic grids of boxes (one of
    them a cell thick), ghosts left NaN and never filled, with each launch
    shape (cells per thread along z, load kind): three steps equal the
    oracle on the whole domain as one box, bit for bit."""
    import torch
    L = O.Lattice(lat)
    dom, lo, n = _boxes(L.D)
    print(f"case direct level lattice={lat} forced={forced} launch={launch} boxes={len(lo)}", flush=True)
    g, ref, per = bc_tables(L, "periodic")
    F = ([2e-5, -1e-5, 3e-6][:L.D]) if forced else None
    omega = 1.0 / 0.7
    steps = 3
    f0 = random_state(L, dom[:L.D], seed=11)
    f = f0.copy()
    for _ in range(steps):
        O.fill_periodic(L, f, dom[:L.D], per)
        out = O.new_fab(L, dom[:L.D])
        O.step(L, f, out, dom[:L.D], omega, F)
        f = out
    expect = O.valid(L, f)
    full0 = O.valid(L, f0)
    ta = [to_dev(_box_fab(L, full0, lo[b], n[b]), dev) for b in range(len(lo))]
    tb = [torch.full_like(t, float("nan")) for t in ta]
    fa = [abi.field(t, n[b][:L.D], 1, L.D) for b, t in enumerate(ta)]
    fb = [abi.field(t, n[b][:L.D], 1, L.D) for b, t in enumerate(tb)]
    lv = abi.Level(abi.LATTICE[lat], dom, per, lo, n)
    try:
        assert lv.direct_available
        lv.set_direct_launch(*launch)
        lv.bind(0, fa)
        lv.bind(1, fb)
        cur = 0
        for _ in range(steps):
            lv.step_direct(cur, 1 - cur, omega, F)
            cur = 1 - cur
        torch.cuda.synchronize()
    finally:
        lv.close()
    final = ta if cur == 0 else tb
    got = _gather(L, dom, lo, n, [to_host(t) for t in final])
    assert np.array_equal(got, expect)
    # the ghosts were never written or needed
    for t in final:
        h = to_host(t)
        assert np.isnan(h).sum() == h.size - np.isfinite(O.valid(L, h)).sum()


@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19"])
def test_direct_level_on_boxes_without_ghosts(dev, lat):
    """Boxes stored without a ghost layer (ngrow 0) serve the direct step,
    which equals the oracle on the whole domain bit for bit; the exchange,
    the boundary fill and the step through ghosts refuse such a slot."""
    import torch
    L = O.Lattice(lat)
    dom, lo, n = _boxes(L.D)
    print(f"case direct level no ghosts lattice={lat}", flush=True)
    g, ref, per = bc_tables(L, "periodic")
    F = [2e-5, -1e-5, 3e-6][:L.D]
    omega = 1.0 / 0.7
    f0 = random_state(L, dom[:L.D], seed=13)
    f = f0.copy()
    for _ in range(3):
        O.fill_periodic(L, f, dom[:L.D], per)
        out = O.new_fab(L, dom[:L.D])
        O.step(L, f, out, dom[:L.D], omega, F)
        f = out
    expect = O.valid(L, f)
    full0 = O.valid(L, f0)
    def box(b):
        l, m = lo[b], n[b]
        if L.D == 3:
            return full0[:, l[2]:l[2] + m[2], l[1]:l[1] + m[1], l[0]:l[0] + m[0]]
        return full0[:, :, l[1]:l[1] + m[1], l[0]:l[0] + m[0]]
    ta = [to_dev(np.ascontiguousarray(box(b)), dev) for b in range(len(lo))]
    tb = [torch.full_like(t, float("nan")) for t in ta]
    lv = abi.Level(abi.LATTICE[lat], dom, per, lo, n)
    try:
        lv.bind(0, [abi.field(t, n[b][:L.D], 0, L.D) for b, t in enumerate(ta)])
        lv.bind(1, [abi.field(t, n[b][:L.D], 0, L.D) for b, t in enumerate(tb)])
        with pytest.raises(abi.MfaError, match="ghost layer"):
            lv.fill_boundary(0)
        with pytest.raises(abi.MfaError, match="ghost layer"):
            lv.apply_bc(0, g)
        with pytest.raises(abi.MfaError, match="ghost layer"):
            lv.step(0, 1, omega, F)
        cur = 0
        for _ in range(3):
            lv.step_direct(cur, 1 - cur, omega, F)
            cur = 1 - cur
        torch.cuda.synchronize()
    finally:
        lv.close()
    final = ta if cur == 0 else tb
    got = np.full((L.Q, dom[2], dom[1], dom[0]), np.nan)
    for b, t in enumerate(final):
        l, m = lo[b], n[b]
        h = to_host(t).reshape(L.Q, m[2], m[1], m[0])
        got[:, l[2]:l[2] + m[2], l[1]:l[1] + m[1], l[0]:l[0] + m[0]] = h
    assert np.array_equal(got, expect.reshape(got.shape))


def test_direct_level_refuses_other_layouts(dev):
    """Boxes that are not a tensor-product grid, or a domain with a wall,
    have no direct step."""
    lv = abi.Level(abi.LATTICE["D2Q9"], [20, 10], [1, 1], [[0, 0], [10, 0], [0, 5], [7, 5]],
                   [[10, 5], [10, 5], [7, 5], [13, 5]])
    try:
        assert not lv.direct_available
        with pytest.raises(abi.MfaError, match="tensor-product grid"):
            lv.step_direct(0, 1, 1.0)
    finally:
        lv.close()
    lv = abi.Level(abi.LATTICE["D2Q9"], [20, 10], [1, 0], [[0, 0], [10, 0]], [[10, 10], [10, 10]])
    try:
        assert not lv.direct_available
    finally:
        lv.close()
    lv = abi.Level(abi.LATTICE["D2Q9"], [20, 10], [1, 1], [[0, 0], [10, 0]], [[10, 10], [10, 10]])
    try:
        assert lv.direct_available
        for bad in ((3, 0), (2, 3)):
            with pytest.raises(abi.MfaError):
                lv.set_direct_launch(*bad)
    finally:
        lv.close()
