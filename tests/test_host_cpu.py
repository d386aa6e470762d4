"""Host (C++) surface on the CPU: header/exports, the inputs parser and the
UnitConverter chain against the reference's own numbers
(tests/unit/test_units.cpp, the lid2d and uniform_force inputs, the
configurations' tau/nu), and the abort conditions of InputParser/Units; and
the benchmark's refusal of more than one GPU."""
import os
import re

import pytest

from muphfasa_amd import host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INP = os.path.join(ROOT, "tests", "golden", "inputs")


def _read(name):
    return open(os.path.join(INP, name)).read()


def test_host_exports_match_header():
    text = open(os.path.join(ROOT, "include", "mfa_host.h")).read()
    declared = set(re.findall(r"^\s*(?:int|void|const char\*|mfa_sim\*)\s+(mfa_\w+)\s*\(", text, re.M))
    assert declared == set(host.EXPORTS)
    L = host.lib()
    for name in declared:
        assert hasattr(L, name)


def test_units_kat_test_units_cpp():
    # tests/unit/test_units.cpp:32-54: L=1, U=1, Re=32, N=32 -> dx=1/32, a_lat = a dt^2/dx
    text = ("grid.ncells = 32 32 32\nsim.max_step = 1\nunits.Re = 32\n"
            "geometry.prob_lo = 0 0 0\ngeometry.prob_hi = 1 1 1\n"
            "bc.lo_type = Periodic Periodic Periodic\nbc.hi_type = Periodic Periodic Periodic\n"
            "forcing.accel = 0.125 0 0\n")
    u = host.inputs_units(text)
    assert abs(u["dx"] - 1.0 / 32.0) <= 1e-15
    a = 0.125
    assert abs(u["F0"] - a * u["dt"] * u["dt"] / u["dx"]) <= 1e-15
    assert abs(u["F0"] - a * u["dx"] * u["u_lat"] ** 2) <= 1e-15
    assert u["F1"] == 0.0


def test_units_of_configurations():
    u = host.inputs_units(_read("lid_cavity_128"))
    assert abs(u["nu_lat"] - 0.128) < 1e-15 and abs(u["tau"] - 0.884) < 1e-15
    u = host.inputs_units(_read("periodic_d3q19_256"))
    assert u["dx"] == 1.0 and u["dt"] == 1.0 and u["omega"] == 1.25 and u["F0"] == 1e-6
    # tests/regression/lid2d/inputs: tau pinned at 0.692 gives u_lat = 0.1 at N=64
    lid2d = _read("lid_cavity_128").replace("128 128", "64 64").replace("lbm.u_lat          = 0.1",
                                                                       "lbm.tau = 0.692")
    lid2d = lid2d.replace("units.ref_velocity = 1.0", "units.ref_velocity = 0.05")
    u = host.inputs_units(lid2d)
    assert abs(u["u_lat"] - 0.1) < 1e-14 and abs(u["tau"] - 0.692) < 1e-14


@pytest.mark.parametrize("override,msg", [
    ("units.nu_phys=0.01", "Both units.Re and units.nu_phys"),
    ("lbm.tau=0.9", "lbm.u_lat and lbm.tau were both"),
    ("lbm.u_lat=0.35", "low Mach"),
    ("geometry.prob_hi=1.0 2.0", "square"),
    ("bc.hiy_Uwall=1.0 0.5", "normal"),
])
def test_input_errors(override, msg):
    with pytest.raises(host.HostError, match=msg):
        host.inputs_units(_read("lid_cavity_128"), [override])


def test_missing_required_key_and_bad_syntax():
    text = _read("lid_cavity_128").replace("grid.ncells        = 128 128\n", "")
    with pytest.raises(host.HostError, match="grid.ncells"):
        host.inputs_units(text)
    with pytest.raises(host.HostError, match="no '='"):
        host.inputs_units("grid.ncells 32 32\n")
    with pytest.raises(host.HostError, match="tau is not within"):
        host.inputs_units(_read("lid_cavity_128"), ["units.Re=0.01"])


def test_parmparse_last_definition_wins_and_comments():
    text = _read("lid_cavity_128") + "\nunits.Re = 400  # later definition wins\n"
    u = host.inputs_units(text)
    assert u["Re"] == 400.0
    u2 = host.inputs_units(text, ["units.Re=100"])     # command-line override wins
    assert u2["Re"] == 100.0


def _inputinfo(*args):
    import subprocess
    exe = os.path.join(ROOT, "muphfasa_amd", "lib", "muphfasa_inputinfo")
    return subprocess.run([exe, *args], capture_output=True, text=True, timeout=30)


def test_inputinfo_summary_matches_unit_converter():
    """The inspector reports the solver's own lattice units and step count."""
    r = _inputinfo(os.path.join(INP, "lid_cavity_128"))
    assert r.returncode == 0, r.stdout + r.stderr
    u = host.inputs_units(_read("lid_cavity_128"))
    tau = float(re.search(r"^\s+tau\s+(\S+)", r.stdout, re.M).group(1))
    dt = float(re.search(r"^\s+dt\s+(\S+) s", r.stdout, re.M).group(1))
    assert abs(tau - u["tau"]) <= 1e-5 * u["tau"] and abs(dt - u["dt"]) <= 1e-5 * u["dt"]
    assert re.search(r"steps\s+10000\s+\(limit: sim.max_step\)", r.stdout)
    assert "lattice (0.1, 0)" in r.stdout and "accepts" in r.stdout


def test_inputinfo_advisories_and_refusals():
    # Re = 10^4 at 128 cells: tau 0.5038, no admissible u_lat: refinement advice, still accepted
    r = _inputinfo(os.path.join(INP, "lid_cavity_128"), "units.Re=10000")
    assert r.returncode == 0 and "[!] tau = 0.5038" in r.stdout and "refine every direction" in r.stdout
    # tau above the ideal range with a feasible fix: the suggested u_lat lands inside it
    r = _inputinfo(os.path.join(INP, "lid_cavity_128"), "units.Re=10", "lbm.u_lat=0.05")
    m = re.search(r"suggest lbm.u_lat = (\S+) \(tau (\S+)\)", r.stdout)
    assert m and 0.6 <= float(m.group(2)) <= 1.5, r.stdout
    # the solver's hard limit (u_lat >= 0.3) and an unknown lattice: exit status 1
    assert _inputinfo(os.path.join(INP, "lid_cavity_128"), "lbm.u_lat=0.35").returncode == 1
    assert _inputinfo(os.path.join(INP, "periodic_d3q19_256"), "lbm.lattice_type=D3Q7").returncode == 1
    r = _inputinfo(os.path.join(INP, "lid_cavity_128"), "units.nu_phys=0.1")
    assert r.returncode == 1 and "Both units.Re and units.nu_phys" in r.stdout


DESC = "((8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5 6 7 8)))"


def _ivect(v):
    return "(" + ",".join(str(int(x)) for x in v) + ")"


def _box(lo, hi):
    return "(" + _ivect(lo) + " " + _ivect(hi) + " " + _ivect([0] * len(lo)) + ")"


def write_reference_checkpoint(d, n, ncomp, f, boxes, ngrow=1, files=2, step=20, time=21.0,
                               version="muphfasa_checkpoint_v1", header_box=None, ngrow_text=None,
                               desc=DESC, fab_grow=None, short=None, skip_file=False, fab=True):
    """A checkpoint directory laid out as the reference's writeCheckpoint
    writes it (src/IO/Checkpoint.cpp:26-59): text Header, then VisMF::Write
    of the populations: f_H listing the valid boxes, the ghost width and each
    box's FAB by file and byte offset; each FAB holds the box grown by the
    ghost width (ghost cells filled with NaN here), ncomp planes, x fastest,
    spread over `files` data files.  f: (ncomp, [n2,] n1, n0) valid cells.
    The keyword arguments plant one defect each."""
    import numpy as np
    os.makedirs(d, exist_ok=True)
    dim = len(n)
    hb = header_box if header_box is not None else _box([0] * dim, [x - 1 for x in n])
    with open(os.path.join(d, "Header"), "w") as h:
        h.write(f"{version}\n{step}\n{time!r}\n{hb}\n{ncomp}\n{_ivect([ngrow] * dim)}\n")
    fg = ngrow if fab_grow is None else fab_grow
    fods, offsets = [], {}
    for i, (lo, hi) in enumerate(boxes):
        name = f"f_D_{i % files:05d}"
        path = os.path.join(d, name)
        glo = [x - fg for x in lo]
        ghi = [x + fg for x in hi]
        ext = [b - a + 1 for a, b in zip(glo, ghi)]
        data = np.full([ncomp] + ext[::-1], np.nan)
        sl = tuple(slice(lo[k] - glo[k], lo[k] - glo[k] + hi[k] - lo[k] + 1) for k in range(dim))[::-1]
        src = tuple(slice(lo[k], hi[k] + 1) for k in range(dim))[::-1]
        data[(slice(None),) + sl] = f[(slice(None),) + src]
        raw = data.astype("<f8").tobytes()
        if short is not None and i == short:
            raw = raw[:-8]
        off = offsets.get(name, 0)
        if not (skip_file and i == len(boxes) - 1):
            with open(path, "ab") as fh:
                head = (f"FAB {desc}{_box(glo, ghi)} {ncomp}\n" if fab else "") .encode()
                fh.write(head + raw)
                offsets[name] = off + len(head) + len(raw)
        fods.append((name, off))
    with open(os.path.join(d, "f_H"), "w") as h:
        h.write(f"1\n0\n{ncomp}\n{ngrow_text if ngrow_text is not None else ngrow}\n")
        h.write(f"({len(boxes)} 0\n")
        for lo, hi in boxes:
            h.write(_box(lo, hi) + "\n")
        h.write(")\n")
        h.write(f"{len(boxes)}\n")
        for name, off in fods:
            h.write(f"FabOnDisk: {name} {off}\n")
        h.write("\n")
    return d


def _fake_checkpoint(d, n, ncomp, **kw):
    """A one-box 2-D checkpoint of n cells with populations 1/9."""
    import numpy as np
    f = np.full((ncomp, n[1], n[0]), 1.0 / 9.0)
    return write_reference_checkpoint(d, n, ncomp, f, [([0, 0], [n[0] - 1, n[1] - 1])], ngrow=0,
                                      files=1, **kw)


def test_inputinfo_validates_restart_checkpoints(tmp_path):
    """amr.restart is checked as restartFromCheckpoint checks it
    (src/IO/Checkpoint.cpp:61-103): a usable checkpoint names the step the
    run resumes at; a missing or malformed Header, a different domain or
    component count, and missing or short data make the inputs refused."""
    inp = os.path.join(INP, "lid_cavity_128")
    n = [16, 16]
    args = ["grid.ncells=16 16"]
    ok = _fake_checkpoint(str(tmp_path / "ok"), n, 9)
    r = _inputinfo(inp, *args, f"amr.restart={ok}")
    assert r.returncode == 0 and "resumes at step 21" in r.stdout, r.stdout
    cases = {
        "no Header file found": lambda d: os.makedirs(d) or d,
        "unrecognized checkpoint version": lambda d: _fake_checkpoint(d, n, 9, version="chk_v0"),
        "malformed Header": lambda d: _fake_checkpoint(d, n, 9, header_box="(0 0 15 15)"),
        "domain mismatch": lambda d: _fake_checkpoint(d, [32, 32], 9),
        "holds 19 distribution components": lambda d: _fake_checkpoint(d, n, 19),
        "does not cover the domain": lambda d: _fake_checkpoint(d, n, 9, short=0),
        "is missing": lambda d: _fake_checkpoint(d, n, 9, skip_file=True),
        "is not a FAB file": lambda d: _fake_checkpoint(d, n, 9, fab=False),
    }
    for i, (msg, make) in enumerate(cases.items()):
        d = make(str(tmp_path / f"bad{i}"))
        r = _inputinfo(inp, *args, f"amr.restart={d}")
        assert r.returncode == 1 and msg in r.stdout and "would refuse" in r.stdout, (msg, r.stdout)


def _checkpoint_read(d, n, ncomp):
    import ctypes
    import numpy as np
    L = host.lib()
    dim = len(n)
    f = np.zeros((ncomp,) + tuple(n[::-1]))
    step, time = ctypes.c_int(), ctypes.c_double()
    nn = (ctypes.c_int * 3)(*(list(n) + [1] * (3 - dim)))
    rc = L.mfa_checkpoint_read(d.encode(), dim, nn, ncomp,
                               f.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                               ctypes.byref(step), ctypes.byref(time))
    return rc, f, step.value, time.value, L.mfa_sim_error().decode()


@pytest.mark.parametrize("dim,n,ncomp,boxes,ngrow,ngrow_text", [
    (3, [8, 6, 5], 19, [([0, 0, 0], [3, 2, 4]), ([4, 0, 0], [7, 2, 4]), ([0, 3, 0], [3, 5, 4]),
                        ([4, 3, 0], [7, 5, 4])], 1, None),
    (2, [13, 7], 9, [([0, 0], [4, 6]), ([5, 0], [12, 2]), ([5, 3], [12, 6])], 1, "(1,1)"),
    (3, [5, 4, 3], 27, [([0, 0, 0], [4, 3, 1]), ([0, 0, 2], [4, 3, 2])], 2, None),
])
def test_restart_reads_reference_checkpoints_of_several_boxes(tmp_path, dim, n, ncomp, boxes, ngrow,
                                                              ngrow_text):
    """A checkpoint in the layout the reference writes (several boxes of
    max_grid_size, each FAB grown by the ghost width, FABs at byte offsets
    in two data files; tests/unit/test_restart.cpp:219-226 restarts across
    box layouts) is read into this engine's one box: every valid cell in its
    place, no ghost value (NaN in the files) taken."""
    import numpy as np
    rng = np.random.default_rng(dim * 100 + ncomp)
    f = rng.standard_normal((ncomp,) + tuple(n[::-1]))
    d = write_reference_checkpoint(str(tmp_path / "chk"), n, ncomp, f, boxes, ngrow=ngrow,
                                   ngrow_text=ngrow_text, step=37, time=38.5)
    rc, got, step, time, err = _checkpoint_read(d, n, ncomp)
    assert rc == 0, err
    assert np.array_equal(got, f)
    assert (step, time) == (37, 38.5)


def test_restart_refuses_malformed_reference_checkpoints(tmp_path):
    """Each defect of the populations' VisMF data is refused with its reason
    (the reference aborts in VisMF::Read or in readCheckpoint's coverage
    check, src/IO/Checkpoint.cpp:131-135), and muphfasa_inputinfo reports the
    same checkpoint unusable."""
    import numpy as np
    n, ncomp = [8, 6, 5], 19
    boxes = [([0, 0, 0], [3, 5, 4]), ([4, 0, 0], [7, 5, 4])]
    f = np.ones((ncomp,) + tuple(n[::-1]))
    cases = {
        "is not its box grown by the ghost width": dict(fab_grow=0),
        "does not cover the domain": dict(short=1),
        "little-endian IEEE doubles": dict(desc=DESC.replace("(1 2 3 4 5 6 7 8)", "(8 7 6 5 4 3 2 1)")),
        "is missing": dict(skip_file=True),
    }
    for i, (msg, kw) in enumerate(cases.items()):
        d = write_reference_checkpoint(str(tmp_path / f"c{i}"), n, ncomp, f, boxes, **kw)
        rc, _, _, _, err = _checkpoint_read(d, n, ncomp)
        assert rc != 0 and msg in err, (msg, err)
    # boxes that overlap or leave a gap
    for i, bx in enumerate([[([0, 0, 0], [4, 5, 4]), ([4, 0, 0], [7, 5, 4])],
                            [([0, 0, 0], [3, 5, 4]), ([5, 0, 0], [7, 5, 4])]]):
        d = write_reference_checkpoint(str(tmp_path / f"b{i}"), n, ncomp, f, bx)
        rc, _, _, _, err = _checkpoint_read(d, n, ncomp)
        assert rc != 0 and ("overlap" in err or "does not cover the domain" in err), err
    inp = os.path.join(INP, "periodic_d3q19_256")
    bad = write_reference_checkpoint(str(tmp_path / "insp"), n, ncomp, f, boxes, fab_grow=0)
    good = write_reference_checkpoint(str(tmp_path / "insp_ok"), n, ncomp, f, boxes)
    args = ["grid.ncells=8 6 5", "geometry.prob_hi=8 6 5"]
    r = _inputinfo(inp, *args, f"amr.restart={good}")
    assert r.returncode == 0 and "resumes at step 21" in r.stdout, r.stdout
    r = _inputinfo(inp, *args, f"amr.restart={bad}")
    assert r.returncode == 1 and "grown by the ghost width" in r.stdout, r.stdout


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus 2 exits with status 2 and a JSON line
    that says so, before it touches a device."""
    import json
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]
