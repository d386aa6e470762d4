"""Builds the C++ host (libmuphfasa_host.so and the muphfasa executable)
against libmfa_lbm.so, in-tree, with hipcc (host code + HIP runtime), and the
host-only inputs inspector muphfasa_inputinfo with g++."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HOST = os.path.join(HERE, "host")
LIBDIR = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["ParmParse.cpp", "InputParser.cpp", "Units.cpp", "PlotIO.cpp", "Simulation.cpp",
           "HostABI.cpp"]
FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-ffp-contract=off", "-I", os.path.join(ROOT, "include")]
# the inputs inspector is host-only (parser + unit converter, no HIP)
INFO_SOURCES = ["input_info.cpp", "ParmParse.cpp", "InputParser.cpp", "Units.cpp", "PlotIO.cpp"]


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    lib = os.path.join(LIBDIR, "libmuphfasa_host.so")
    exe = os.path.join(LIBDIR, "muphfasa")
    srcs = [os.path.join(HOST, s) for s in SOURCES]
    link = ["-L", LIBDIR, "-lmfa_lbm", "-Wl,-rpath,$ORIGIN"]
    cmds = [
        [HIPCC, *FLAGS, "-shared", "-o", lib + ".tmp", *srcs, *link],
        [HIPCC, *FLAGS, "-o", exe + ".tmp", os.path.join(HOST, "main.cpp"), "-L", LIBDIR,
         "-lmuphfasa_host", "-lmfa_lbm", "-Wl,-rpath,$ORIGIN"],
    ]
    info = os.path.join(LIBDIR, "muphfasa_inputinfo")
    cmds.append(["g++", "-O2", "-std=c++17", "-Wall", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                 "-o", info + ".tmp", *[os.path.join(HOST, s) for s in INFO_SOURCES]])
    for i, cmd in enumerate(cmds):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        if i == 0:
            os.replace(lib + ".tmp", lib)
    os.replace(exe + ".tmp", exe)
    os.replace(info + ".tmp", info)
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
