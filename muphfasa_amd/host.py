"""ctypes binding of the C++ host (include/mfa_host.h): the simulation API
(ISimulation/makeSimulation) and the solver run loop, for Python drivers and
the test-suite."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmuphfasa_host.so")
EXE_PATH = os.path.join(HERE, "lib", "muphfasa")
EXPORTS = ["mfa_sim_error", "mfa_sim_create", "mfa_sim_destroy", "mfa_sim_initialize",
           "mfa_sim_advance", "mfa_sim_update_macroscopic", "mfa_sim_write_output", "mfa_sim_sync",
           "mfa_sim_info", "mfa_sim_copy_f", "mfa_sim_copy_macro", "mfa_sim_health", "mfa_inputs_units",
           "mfa_run_main", "mfa_checkpoint_read", "mfa_sim_stage_times"]
STAGES = ["advance", "updateMacroscopic", "writePlotfile", "writeCheckpoint", "initialize"]
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build with `python -m muphfasa_amd.host_build`")
        L = abi.load_library(LIB_PATH)
        cp = ctypes.c_char_p
        L.mfa_sim_error.restype = cp
        L.mfa_sim_create.restype = ctypes.c_void_p
        L.mfa_sim_create.argtypes = [cp, ctypes.c_int, ctypes.POINTER(cp)]
        L.mfa_sim_destroy.argtypes = [ctypes.c_void_p]
        for name in ("mfa_sim_initialize", "mfa_sim_sync"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        for name in ("mfa_sim_advance", "mfa_sim_update_macroscopic", "mfa_sim_write_output"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_int]
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.mfa_sim_info.argtypes = [ctypes.c_void_p, ip, ip, dp, dp, dp, ip]
        L.mfa_sim_copy_f.argtypes = [ctypes.c_void_p, dp]
        L.mfa_sim_copy_macro.argtypes = [ctypes.c_void_p, dp, dp]
        L.mfa_sim_health.argtypes = [ctypes.c_void_p, dp]
        L.mfa_inputs_units.argtypes = [cp, ctypes.c_int, ctypes.POINTER(cp), dp]
        L.mfa_run_main.argtypes = [ctypes.c_int, ctypes.POINTER(cp)]
        L.mfa_checkpoint_read.argtypes = [cp, ctypes.c_int, ip, ctypes.c_int, dp, ip, dp]
        L.mfa_sim_stage_times.argtypes = [ctypes.c_void_p, dp, ctypes.POINTER(ctypes.c_longlong)]
        _lib = L
    return _lib


class HostError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise HostError(lib().mfa_sim_error().decode())


def _strs(items):
    arr = (ctypes.c_char_p * max(1, len(items)))()
    for i, s in enumerate(items):
        arr[i] = s.encode()
    return arr


UNIT_KEYS = ["dx", "dt", "u_lat", "nu_lat", "tau", "omega", "Re", "nu_phys", "Ma", "F0", "F1", "F2"]


def inputs_units(text: str, overrides=()):
    out = (ctypes.c_double * 12)()
    _check(lib().mfa_inputs_units(text.encode(), len(overrides), _strs(list(overrides)), out))
    return dict(zip(UNIT_KEYS, list(out)))


class Simulation:
    """ISimulation through the host C ABI."""

    def __init__(self, inputs_text: str, overrides=()):
        self._h = lib().mfa_sim_create(inputs_text.encode(), len(overrides), _strs(list(overrides)))
        if not self._h:
            raise HostError(lib().mfa_sim_error().decode())
        lat = ctypes.c_int()
        n = (ctypes.c_int * 3)()
        self._info = (lat, n)
        self.refresh()

    def refresh(self):
        lat, n = self._info
        om = ctypes.c_double()
        F = (ctypes.c_double * 3)()
        t = ctypes.c_double()
        s0 = ctypes.c_int()
        _check(lib().mfa_sim_info(self._h, ctypes.byref(lat), n, ctypes.byref(om), F,
                                  ctypes.byref(t), ctypes.byref(s0)))
        self.lattice = lat.value
        self.n = list(n)
        self.omega = om.value
        self.force = list(F)
        self._time = t.value
        self._start_step = s0.value
        self.Q = {0: 9, 1: 15, 2: 19, 3: 27}[self.lattice]
        self.D = 2 if self.lattice == 0 else 3

    @property
    def time(self):
        self.refresh()
        return self._time

    @property
    def start_step(self):
        self.refresh()
        return self._start_step

    def initialize(self):
        _check(lib().mfa_sim_initialize(self._h))

    def advance(self, nsteps=1):
        _check(lib().mfa_sim_advance(self._h, nsteps))

    def update_macroscopic(self, step):
        _check(lib().mfa_sim_update_macroscopic(self._h, step))

    def write_output(self, step):
        _check(lib().mfa_sim_write_output(self._h, step))

    def sync(self):
        _check(lib().mfa_sim_sync(self._h))

    def f(self):
        a = np.empty((self.Q, self.n[2], self.n[1], self.n[0]))
        _check(lib().mfa_sim_copy_f(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return a

    def macro(self):
        r = np.empty((self.n[2], self.n[1], self.n[0]))
        v = np.empty((self.D, self.n[2], self.n[1], self.n[0]))
        dp = ctypes.POINTER(ctypes.c_double)
        _check(lib().mfa_sim_copy_macro(self._h, r.ctypes.data_as(dp), v.ctypes.data_as(dp)))
        return r, v

    def stage_times(self):
        """{stage: (total ms, calls)} since creation (mfa_sim_stage_times)."""
        ms = (ctypes.c_double * 5)()
        calls = (ctypes.c_longlong * 5)()
        _check(lib().mfa_sim_stage_times(self._h, ms, calls))
        return {name: (ms[i], calls[i]) for i, name in enumerate(STAGES)}

    def health(self):
        """(cells with a non-finite population, cells with finite density <= 0)."""
        out = (ctypes.c_double * 2)()
        _check(lib().mfa_sim_health(self._h, out))
        return int(out[0]), int(out[1])

    def close(self):
        if self._h:
            lib().mfa_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_plotfile(path):
    """Reader for the single-level plotfiles the host writes: returns
    (names, time, step, data[ncomp, n2, n1, n0])."""
    with open(os.path.join(path, "Header")) as fh:
        lines = fh.read().split("\n")
    assert lines[0] == "HyperCLaw-V1.1"
    nv = int(lines[1])
    names = lines[2:2 + nv]
    dim = int(lines[2 + nv])
    time = float(lines[3 + nv])
    box = lines[8 + nv].strip()
    step = int(lines[9 + nv].split()[0])
    nums = [int(t) for t in box.replace("(", " ").replace(")", " ").replace(",", " ").split()]
    hi = nums[dim:2 * dim]
    n = [h + 1 for h in hi] + [1] * (3 - dim)
    with open(os.path.join(path, "Level_0", "Cell_D_00000"), "rb") as fh:
        hdr = fh.readline()
        assert hdr.startswith(b"FAB ")
        data = np.frombuffer(fh.read(), dtype="<f8")
    return names, time, step, data.reshape(nv, n[2], n[1], n[0])
