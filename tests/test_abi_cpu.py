"""CPU-side checks of the C ABI: the library loads, exports every symbol the
header declares, validates arguments before touching the GPU, and its host
boundary resolution (mfa_resolve_bc) equals the oracle's BCManager
restatement for every face combination."""
import ctypes
import itertools
import os
import re

import numpy as np
import pytest

from muphfasa_amd import abi
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_match_header():
    """abi.EXPORTS lists exactly the functions of mfa_lbm.h and mfa_level.h."""
    declared = set()
    for h in ("mfa_lbm.h", "mfa_level.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        declared |= set(re.findall(r"^\s*(?:int|long long|const char\*)\s+(mfa_\w+)\s*\(", text, re.M))
    assert declared, "no declarations parsed"
    L = abi.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert declared == set(abi.EXPORTS)


def test_version():
    assert b"gfx950" in abi.lib().mfa_version()


NAMES = ["MovingWall", "NoSlip", "Inflow", "Outflow", "PressureOutlet"]


@pytest.mark.parametrize("lattice", ["D2Q9", "D3Q19"])
def test_resolve_bc_matches_oracle(lattice):
    lid = abi.LATTICE[lattice]
    D = abi.LATTICE_DQ[lid][0]
    rng = np.random.default_rng(5)
    choices = NAMES + ["Periodic"]
    combos = 0
    for trial in range(400):
        ft = []
        for d in range(D):
            if rng.random() < 0.25:
                ft += [abi.BC["Periodic"]] * 2
            else:
                ft += [abi.BC[rng.choice(NAMES)], abi.BC[rng.choice(NAMES)]]
        fr = rng.uniform(0.9, 1.1, 6)
        fu = rng.uniform(-0.1, 0.1, (6, 3))
        try:
            ref = O.resolve_bc(D, ft, fr, fu)
        except ValueError:
            with pytest.raises(abi.MfaError) as e:
                abi.resolve_bc(lid, ft, fr, fu)
            assert e.value.code == -4
            continue
        g = abi.resolve_bc(lid, ft, fr, fu)
        assert list(g.resolved_type) == list(ref[0])
        assert np.array_equal(np.array(g.resolved_rho), ref[1])
        assert np.array_equal(np.array([list(r) for r in g.resolved_u]), ref[2])
        assert list(g.periodic)[:D] == [int(ft[2 * d] == abi.BC["Periodic"]) for d in range(D)]
        combos += 1
    assert combos > 50


def test_invalid_arguments_fail_before_launch():
    g = abi.resolve_bc(2, [abi.BC["Periodic"]] * 6)
    f = abi.Field()
    with pytest.raises(abi.MfaError) as e:
        abi.check(abi.lib().mfa_step(ctypes.byref(f), ctypes.byref(f), ctypes.byref(g), 1.0,
                                     None, None))
    assert e.value.code == -1
    # a Periodic hi face on a non-periodic direction (lo face not Periodic) is rejected
    with pytest.raises(abi.MfaError):
        abi.resolve_bc(2, [abi.BC["NoSlip"], abi.BC["Periodic"]] + [abi.BC["Periodic"]] * 4)
    # a Periodic lo face makes the direction periodic whatever the hi face says
    g = abi.resolve_bc(2, [abi.BC["Periodic"], abi.BC["NoSlip"]] + [abi.BC["Periodic"]] * 4)
    assert list(g.periodic) == [1, 1, 1]
    # unsupported dtype is reported, not silently run
    f.data = 16
    f.n[0] = f.n[1] = f.n[2] = 4
    f.sy, f.sz, f.sq, f.ncomp, f.dtype = 4, 16, 64, 19, 3
    with pytest.raises(abi.MfaError) as e:
        abi.check(abi.lib().mfa_pull_collide(ctypes.byref(f), ctypes.byref(f), 2, 1.0, None,
                                             None))
    assert e.value.code in (-1, -2)


def test_launch_configuration_tables():
    """mfa_set_launch / mfa_set_launch_dev validate, store and reset (host-only)."""
    lat = abi.LATTICE["D3Q27"]
    default = abi.get_launch(lat)
    abi.set_launch(lat, 128, 2, 1, 1)
    assert abi.get_launch(lat) == (128, 2, 1, 1)
    for bad in ((96, 1, 1, 0), (256, 2, 1, 0), (64, 1, 3, 0), (64, 1, 1, 3)):
        with pytest.raises(abi.MfaError):
            abi.set_launch(lat, *bad)
    abi.set_launch(lat)
    assert abi.get_launch(lat) == default
    for dt in ("f32", "f16dev", "bf16dev"):
        d = abi.get_launch_dev(lat, abi.DTYPE[dt])
        abi.set_launch_dev(lat, abi.DTYPE[dt], 128, 2, 1, 0)
        assert abi.get_launch_dev(lat, abi.DTYPE[dt]) == (128, 2, 1, 0)
        with pytest.raises(abi.MfaError):
            abi.set_launch_dev(lat, abi.DTYPE[dt], 32, 1, 2, 1)
        abi.set_launch_dev(lat, abi.DTYPE[dt])
        assert abi.get_launch_dev(lat, abi.DTYPE[dt]) == d
    with pytest.raises(abi.MfaError):
        abi.get_launch_dev(lat, abi.DTYPE["f64"])


def test_overlapping_strides_are_rejected_before_launch():
    """The stride check admits any non-overlapping order (component-major or
    row-interleaved) and rejects views whose valid cells would alias; the
    rejection happens before any device work."""
    g = abi.resolve_bc(2, [abi.BC["Periodic"]] * 6)
    n = 8
    for sy, sz, sq in ((n, n * n, n * n * n - 1),           # component planes overlap
                       (19 * n - 1, 19 * n * n, n),         # rows overlap (row-interleaved)
                       (n, n * n, n // 2),                  # population rows overlap
                       (n - 1, n * n, n * n * n)):          # y rows overlap
        f = abi.Field()
        f.data = 256
        f.n[0] = f.n[1] = f.n[2] = n
        f.sy, f.sz, f.sq, f.ncomp, f.dtype = sy, sz, sq, 19, 0
        with pytest.raises(abi.MfaError) as e:
            abi.check(abi.lib().mfa_step(ctypes.byref(f), ctypes.byref(f), ctypes.byref(g), 1.0,
                                         None, None))
        assert e.value.code == -1 and "strides" in str(e.value), (sy, sz, sq)


def test_every_header_declaration_is_exported():
    """Every function declared in include/*.h is exported by the built
    library that implements it (mfa_host.h: the C++ host; the others: the HIP
    engine).  Only loads and looks up symbols; nothing runs on a device."""
    decl = re.compile(r"^\s*(?:const\s+)?[A-Za-z_]\w*\s*\*?\s*\*?\s*((?:mfa|muphfasa)_\w+)\s*\(",
                      re.M)
    libs = {"mfa_host.h": os.path.join(ROOT, "muphfasa_amd", "lib", "libmuphfasa_host.so")}
    seen = 0
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        names = set(decl.findall(open(os.path.join(ROOT, "include", h)).read()))
        assert names, h
        L = abi.load_library(libs[h]) if h in libs else abi.lib()
        missing = [n for n in sorted(names) if not hasattr(L, n)]
        assert not missing, (h, missing)
        seen += len(names)
    assert seen >= 50


def test_level_layout_checks_need_no_gpu():
    """mfa_level_create validates the box layout on the host before it
    touches the device: overlapping boxes, gaps, boxes outside the domain and
    2-D boxes with a z extent are refused with MFA_ERR_INVALID."""
    from muphfasa_amd import abi
    D2 = abi.LATTICE["D2Q9"]
    bad = [
        ([8, 8], [[0, 0], [4, 0]], [[5, 8], [4, 8]]),      # overlap
        ([8, 8], [[0, 0]], [[8, 7]]),                      # gap
        ([8, 8], [[0, 0], [4, 0]], [[4, 8], [5, 8]]),      # past the domain
        ([8, 8], [[0, 0, 1]], [[8, 8, 1]]),                # 2-D box off z = 0
    ]
    for dom, lo, n in bad:
        with pytest.raises(abi.MfaError) as e:
            abi.Level(D2, dom, [1, 1], lo, n)
        assert e.value.code == -1, (dom, lo, n)


def test_level_chopping_tiles_the_domain():
    """muphfasa_amd.level.chop / boxes: the max_grid_size layout covers the
    domain exactly, every box is at most max_grid cells per side, and the
    pieces of one dimension differ by at most one cell."""
    from muphfasa_amd import level
    for n, mg in ((256, 64), (100, 64), (7, 64), (130, 32), (1, 1)):
        pieces = level.chop(n, mg)
        assert sum(m for _, m in pieces) == n
        assert all(m <= mg for _, m in pieces)
        assert max(m for _, m in pieces) - min(m for _, m in pieces) <= 1
        assert [lo for lo, _ in pieces] == [sum(m for _, m in pieces[:i]) for i in range(len(pieces))]
    lo, n = level.boxes([40, 30, 20], 16)
    cover = set()
    for l, m in zip(lo, n):
        for z in range(l[2], l[2] + m[2]):
            for y in range(l[1], l[1] + m[1]):
                for x in range(l[0], l[0] + m[0]):
                    assert (x, y, z) not in cover
                    cover.add((x, y, z))
    assert len(cover) == 40 * 30 * 20


def test_overlapping_input_and_output_are_rejected_before_launch():
    """mfa_step, mfa_step2, mfa_pull_collide and mfa_convert read one field
    while writing another: views whose byte ranges (all components, ghost
    layers included) intersect are refused with MFA_ERR_INVALID before any
    device work; views that touch end to end pass this check (shown by the
    next check refusing them for another reason)."""
    g = abi.resolve_bc(abi.LATTICE["D3Q19"], [abi.BC["Periodic"]] * 6)
    n = 8

    def fab(addr, ngrow=0, dtype=0):
        f = abi.Field()
        f.data = addr
        e = n + 2 * ngrow
        f.n[0] = f.n[1] = f.n[2] = n
        f.ngrow = ngrow
        f.sy, f.sz, f.sq, f.ncomp, f.dtype = e, e * e, e * e * e, 19, dtype
        return f

    L = abi.lib()
    size = 19 * n ** 3 * 8
    base = 1 << 30
    for shift in (0, 8, size - 8):                      # same start, one cell on, last element shared
        a, b = fab(base), fab(base + shift)
        for call in (lambda: L.mfa_step(ctypes.byref(a), ctypes.byref(b), ctypes.byref(g), 1.0, None, None),
                     lambda: L.mfa_step2(ctypes.byref(a), ctypes.byref(b), ctypes.byref(g), 1.0, None, None)):
            with pytest.raises(abi.MfaError) as e:
                abi.check(call())
            assert e.value.code == -1 and "overlap" in str(e.value), shift
    # ghost layers count: an FP32 output [base, base + size32); an FP64
    # ghosted input whose lowest ghost byte lies in the output's last element
    # overlaps, one that starts right after it does not (and is then refused
    # only because the storage types differ)
    e1 = n + 2
    reach = 8 * (1 + e1 + e1 * e1)                     # bytes from the lowest ghost to element (0,0,0)
    size32 = size // 2
    out = fab(base, dtype=abi.DTYPE["f32"])
    for extra, refused in ((0, True), (4, False)):
        b = fab(base + size32 - 4 + reach + extra, ngrow=1)
        with pytest.raises(abi.MfaError) as e:
            abi.check(L.mfa_step(ctypes.byref(b), ctypes.byref(out), ctypes.byref(g), 1.0, None, None))
        assert ("overlap" in str(e.value)) == refused, (extra, str(e.value))
        assert refused or "storage types differ" in str(e.value)
    # FP32 output inside an FP64 input's range
    a, b = fab(base), fab(base + size // 2, dtype=abi.DTYPE["f32"])
    with pytest.raises(abi.MfaError) as e:
        abi.check(L.mfa_convert(ctypes.byref(a), ctypes.byref(b), abi.LATTICE["D3Q19"], None))
    assert "overlap" in str(e.value)


def test_pull_collide_refuses_an_output_in_the_z_ghost_of_a_one_plane_box():
    """A 3-D lattice's halo-mode step reads the z ghost planes even of a box
    one plane thick, so an output placed right after the input's last
    component but inside its z = +1 ghost plane overlaps the input.  The call
    is refused on the host, before any GPU work (the pointers are never
    dereferenced)."""
    Q, n = 19, [4, 4, 1]
    e = [n[0] + 2, n[1] + 2, n[2] + 2]
    fin = abi.Field()
    fin.n[0], fin.n[1], fin.n[2] = n
    fin.ngrow = 1
    fin.sy, fin.sz, fin.sq = e[0], e[0] * e[1], e[0] * e[1] * e[2]
    fin.ncomp, fin.dtype = Q, 0
    base = 1 << 40
    fin.data = base
    # last valid element of the last component, then its z = +1 ghost plane
    last_valid = (n[0] - 1) + (n[1] - 1) * fin.sy + (n[2] - 1) * fin.sz + (Q - 1) * fin.sq
    fout = abi.Field()
    fout.n[0], fout.n[1], fout.n[2] = n
    fout.ngrow = 0
    fout.sy, fout.sz, fout.sq = n[0], n[0] * n[1], n[0] * n[1] * n[2]
    fout.ncomp, fout.dtype = Q, 0
    fout.data = base + 8 * (last_valid + 1 + 1 + fin.sy)    # beyond the x/y ghosts of that plane
    rc = abi.lib().mfa_pull_collide(ctypes.byref(fin), ctypes.byref(fout), abi.LATTICE["D3Q19"], 1.0,
                                    None, None)
    assert rc == -1 and b"overlap" in abi.lib().mfa_last_error()


def test_step2_refuses_planes_beyond_32_bit_offsets_before_launch():
    """The 3-D pair kernels address a z-plane through 32-bit byte offsets from
    a wave-uniform plane base, so mfa_step2 refuses (MFA_ERR_UNSUPPORTED, no
    device work; the pointers are never dereferenced) a box whose plane spans
    more than 4 GiB, in FP64 and FP32 storage alike.  ny_fit is the tallest
    plane whose last in-plane byte lies below 2^32 (FP64: 512 + 55188 * 9728
    elements = 4 294 959 104 bytes); one row more crosses it."""
    g = abi.resolve_bc(abi.LATTICE["D3Q19"], [abi.BC["Periodic"]] * 6)
    L = abi.lib()

    def fab(addr, nx, ny, es, dtype):
        f = abi.Field()
        f.data = addr
        f.n[0], f.n[1], f.n[2] = nx, ny, 2
        f.ngrow = 0
        f.sy, f.sq = 19 * nx, nx             # row-interleaved order
        f.sz, f.ncomp, f.dtype = 19 * nx * ny, 19, dtype
        return f

    for es, dtype in ((8, abi.DTYPE["f64"]), (4, abi.DTYPE["f32"])):
        nx = 512
        # in-plane elements of the last cell + 1 = nx + (ny - 1) * 19 nx
        ny_fit = (2 ** 32 // es - nx) // (19 * nx) + 1
        assert (nx + (ny_fit - 1) * 19 * nx) * es <= 2 ** 32 - 1
        assert (nx + ny_fit * 19 * nx) * es > 2 ** 32 - 1
        a = fab(1 << 40, nx, ny_fit + 1, es, dtype)
        b = fab(1 << 44, nx, ny_fit + 1, es, dtype)
        rc = L.mfa_step2(ctypes.byref(a), ctypes.byref(b), ctypes.byref(g), 1.25, None, None)
        msg = L.mfa_last_error().decode()
        assert rc == -2 and "4 GiB" in msg, (es, rc, msg)


def test_convolutions_refuse_misaligned_tensors_before_launch():
    """The warm-start network's convolutions move 16 bytes per lane, so a
    tensor that is not 16-byte aligned is refused (-1) on the host; the
    pointers are never dereferenced."""
    from muphfasa_amd import surrogate as S
    L = S.lib()
    ok, bad = 1 << 20, (1 << 20) + 2
    shape = (1, 8, 8, 16, 16, 1)
    assert L.mfa_nn_conv3x3(bad, ok, None, ok, None, *shape, 1, None) == -1
    assert "aligned" in L.mfa_nn_error().decode()
    assert L.mfa_nn_conv3x3_bwd_data(ok, ok, bad, None, *shape, 0, None) == -1
    assert "aligned" in L.mfa_nn_error().decode()
    assert L.mfa_nn_conv3x3_bwd_weight(ok, bad, ok, None, *shape, None) == -1
    assert "aligned" in L.mfa_nn_error().decode()
