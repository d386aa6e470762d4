"""CPU builds with AddressSanitizer and UndefinedBehaviorSanitizer: the C
oracle (driven over every lattice and boundary kind by
tests/sanitize_oracle_driver.c) and the C++ host's inputs parser and unit
converter (the muphfasa_inputinfo tool run on every inputs file of the tests
and on malformed inputs).  A sanitizer report, or a crash, fails the test."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")


def _clean(res):
    text = res.stderr + res.stdout
    assert "AddressSanitizer" not in text and "runtime error:" not in text and "LeakSanitizer" not in text, text[-3000:]


@pytest.fixture(scope="module")
def builds(tmp_path_factory):
    if not shutil.which("gcc") or not shutil.which("g++"):
        pytest.skip("gcc/g++ not available")
    d = tmp_path_factory.mktemp("san")
    oracle = d / "oracle_san"
    subprocess.run(["gcc", *SAN, "-std=c11", "-ffp-contract=off", "-fopenmp", "-o", str(oracle),
                    os.path.join(ROOT, "tests", "sanitize_oracle_driver.c"),
                    os.path.join(ROOT, "oracle", "lbm_oracle.c"), "-lm"], check=True)
    host = os.path.join(ROOT, "muphfasa_amd", "host")
    info = d / "inputinfo_san"
    subprocess.run(["g++", *SAN, "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                    "-o", str(info)] + [os.path.join(host, f) for f in
                                        ("input_info.cpp", "ParmParse.cpp", "InputParser.cpp", "Units.cpp", "PlotIO.cpp")],
                   check=True)
    return oracle, info


def test_oracle_under_sanitizers(builds):
    res = subprocess.run([str(builds[0])], capture_output=True, text=True, env=dict(ENV, OMP_NUM_THREADS="2"),
                         timeout=300)
    _clean(res)
    assert res.returncode == 0, res.stderr[-2000:]
    assert "ok" in res.stdout


def test_inputs_parser_under_sanitizers(builds, tmp_path):
    inputs = os.path.join(ROOT, "tests", "golden", "inputs")
    cases = [[os.path.join(inputs, f)] for f in sorted(os.listdir(inputs))]
    bad = tmp_path / "bad_inputs"
    bad.write_text("grid.ncells = 12 x\nlbm.lattice_type = D5Q99\nlbm.tau = 0.2\nbc.lo_type = Foo\n"
                   "units.Re = \n# comment only\nunterminated = \"abc\n")
    cases.append([str(bad)])
    cases.append([os.path.join(inputs, "lid_cavity_128"), "grid.ncells=64 64", "lbm.u_lat=0.05",
                  "sim.max_step=10"])
    cases.append([os.path.join(inputs, "lid_cavity_128"), "lbm.tau=0.51", "bc.hi_type=Inflow Outflow"])
    for args in cases:
        res = subprocess.run([str(builds[1]), *args], capture_output=True, text=True, env=ENV, timeout=120)
        _clean(res)
        assert res.returncode in (0, 1), (args, res.returncode, res.stderr[-2000:])
