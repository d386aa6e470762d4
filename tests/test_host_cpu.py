"""Host (C++) surface on the CPU: header/exports, the inputs parser and the
UnitConverter chain against the reference's own numbers
(tests/unit/test_units.cpp, the lid2d and uniform_force inputs, the
configurations' tau/nu), and the abort conditions of InputParser/Units."""
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


def _fake_checkpoint(d, n, ncomp, step=20, time=21.0, version="muphfasa_checkpoint_v1", box=None,
                     data_doubles=None, fab=True):
    """A checkpoint directory in the muphfasa_checkpoint_v1 layout (text
    Header; VisMF f_H / f_D_00000 with one FAB) for a 2-D box of n cells."""
    import numpy as np
    os.makedirs(d, exist_ok=True)
    hi = f"({n[0] - 1},{n[1] - 1})"
    box = box if box is not None else f"((0,0) {hi} (0,0))"
    with open(os.path.join(d, "Header"), "w") as h:
        h.write(f"{version}\n{step}\n{time!r}\n{box}\n{ncomp}\n(0,0)\n")
    if data_doubles is None:
        data_doubles = n[0] * n[1] * ncomp
    if data_doubles >= 0:
        with open(os.path.join(d, "f_D_00000"), "wb") as f:
            if fab:
                f.write(f"FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5 6 7 8)))((0,0) {hi} (0,0)) "
                        f"{ncomp}\n".encode())
            f.write(np.full(data_doubles, 1.0 / 9.0).tobytes())
    return d


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
        "malformed Header": lambda d: _fake_checkpoint(d, n, 9, box="(0 0 15 15)"),
        "domain mismatch": lambda d: _fake_checkpoint(d, [32, 32], 9),
        "holds 19 distribution components": lambda d: _fake_checkpoint(d, n, 19),
        "does not cover the domain": lambda d: _fake_checkpoint(d, n, 9, data_doubles=16 * 16 * 9 - 1),
        "is missing": lambda d: _fake_checkpoint(d, n, 9, data_doubles=-1),
        "is not a FAB file": lambda d: _fake_checkpoint(d, n, 9, fab=False),
    }
    for i, (msg, make) in enumerate(cases.items()):
        d = make(str(tmp_path / f"bad{i}"))
        r = _inputinfo(inp, *args, f"amr.restart={d}")
        assert r.returncode == 1 and msg in r.stdout and "would refuse" in r.stdout, (msg, r.stdout)
