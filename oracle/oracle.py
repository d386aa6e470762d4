"""ctypes wrapper around the CPU oracle (oracle/lbm_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's cpu_baseline leg, always as the checker, never as the
thing measured or shipped.  The product package ``muphfasa_amd`` never imports
this module.

Arrays follow the oracle's ghosted-fab convention: a C-order float64 array of
shape (Q, e2, e1, e0) with e_d = n_d + 2 for the lattice's D dimensions and
e2 = 1 for 2-D lattices.  ``valid(f)`` returns the (Q, n2, n1, n0) view of the
valid cells.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liblbm_oracle.so")

LATTICES = {"D2Q9": 0, "D3Q15": 1, "D3Q19": 2, "D3Q27": 3}
BC = {"MovingWall": 0, "NoSlip": 1, "Inflow": 2, "Outflow": 3, "PressureOutlet": 4,
      "Periodic": 5}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build() -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def use_library(path: str) -> None:
    """Load the oracle from another build of lbm_oracle.c (the CPU baseline
    builds one with -march=native on the host it times)."""
    global _LIB_PATH, _lib
    _LIB_PATH = path
    _lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_feq.restype = ctypes.c_double
        L.orc_feq.argtypes = [ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int]
        L.orc_guo.restype = ctypes.c_double
        L.orc_guo.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_double]
        L.orc_bgk.restype = ctypes.c_double
        L.orc_bgk.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_cell_moments.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_bb_contrib.restype = ctypes.c_double
        L.orc_bb_contrib.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp]
        L.orc_resolve_bc.restype = ctypes.c_int
        L.orc_resolve_bc.argtypes = [ctypes.c_int, _ip, _dp, _dp, _ip, _dp, _dp]
        L.orc_has_conflict.restype = ctypes.c_int
        L.orc_has_conflict.argtypes = [ctypes.c_int, _ip]
        L.orc_step.argtypes = [ctypes.c_int, _dp, _dp, _ip, ctypes.c_double, _dp,
                               ctypes.c_int]
        L.orc_fill_periodic.argtypes = [ctypes.c_int, _dp, _ip, _ip]
        L.orc_step_masked.argtypes = [ctypes.c_int, _dp, _dp, _ip, ctypes.c_double,
                                      ctypes.c_void_p]
        L.orc_apply_bc.restype = ctypes.c_int
        L.orc_apply_bc.argtypes = [ctypes.c_int, _dp, _ip, _ip, _ip, _dp, _dp]
        L.orc_stream.argtypes = [ctypes.c_int, _dp, _dp, _ip]
        L.orc_moments.argtypes = [ctypes.c_int, _dp, _ip, _dp, _dp, _dp, ctypes.c_int]
        L.orc_init_equilibrium.argtypes = [ctypes.c_int, _dp, _dp, _dp, _ip, _dp]
        L.orc_run.restype = ctypes.c_int
        L.orc_run.argtypes = [ctypes.c_int, _dp, _dp, _ip, ctypes.c_double, _dp, _ip,
                              _ip, _dp, _dp, ctypes.c_int, ctypes.c_int]
        L.orc_time_run.restype = ctypes.c_double
        L.orc_time_run.argtypes = L.orc_run.argtypes
        _lib = L
    return _lib


def _d(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _i(seq):
    arr = (ctypes.c_int * len(seq))(*[int(v) for v in seq])
    return arr


def _force(F):
    if F is None:
        return None, None
    a = np.zeros(3, dtype=np.float64)
    a[: len(F)] = F
    return a, _d(a)


class _OrcLattice(ctypes.Structure):
    _fields_ = [("D", ctypes.c_int), ("Q", ctypes.c_int), ("cx", ctypes.c_int * 27),
                ("cy", ctypes.c_int * 27), ("cz", ctypes.c_int * 27),
                ("w", ctypes.c_double * 27), ("opp", ctypes.c_int * 27)]


class Lattice:
    def __init__(self, name: str):
        self.name = name
        self.id = LATTICES[name]
        p = lib().orc_get_lattice
        p.restype = ctypes.POINTER(_OrcLattice)
        p.argtypes = [ctypes.c_int]
        s = p(self.id).contents
        self.D, self.Q = s.D, s.Q
        self.c = np.array([[s.cx[q], s.cy[q], s.cz[q]] for q in range(self.Q)],
                          dtype=np.int64)
        self.w = np.array([s.w[q] for q in range(self.Q)])
        self.opp = np.array([s.opp[q] for q in range(self.Q)], dtype=np.int64)


def fab_shape(lat: Lattice, n):
    n = list(n) + [1] * (3 - len(n))
    e = [n[d] + 2 if d < lat.D else n[d] for d in range(3)]
    return (lat.Q, e[2], e[1], e[0])


def new_fab(lat: Lattice, n, fill=np.nan):
    return np.full(fab_shape(lat, n), fill, dtype=np.float64)


def valid(lat: Lattice, f):
    if lat.D == 2:
        return f[:, :, 1:-1, 1:-1]
    return f[:, 1:-1, 1:-1, 1:-1]


def n3(n):
    n = list(n) + [1] * (3 - len(n))
    return n


def feq(lat, rho, u, q):
    u3 = np.zeros(3)
    u3[: len(u)] = u
    return lib().orc_feq(lat.id, rho, _d(u3), q)


def guo(lat, q, u, F, omega):
    u3 = np.zeros(3)
    u3[: len(u)] = u
    F3 = np.zeros(3)
    F3[: len(F)] = F
    return lib().orc_guo(lat.id, q, _d(u3), _d(F3), omega)


def bgk(f, fe, omega):
    return lib().orc_bgk(f, fe, omega)


def cell_moments(lat, f, F=None):
    f = np.ascontiguousarray(f, dtype=np.float64)
    rho = np.zeros(1)
    u = np.zeros(3)
    Fa, Fp = _force(F)
    lib().orc_cell_moments(lat.id, _d(f), Fp, _d(rho), _d(u))
    return rho[0], u[: lat.D].copy()


def bb_contrib(lat, m, rho, U):
    U3 = np.zeros(3)
    U3[: len(U)] = U
    return lib().orc_bb_contrib(lat.id, m, rho, _d(U3))


def has_conflict(types):
    return bool(lib().orc_has_conflict(len(types), _i(types)))


def resolve_bc(dim, face_type, face_rho=None, face_u=None):
    """BCManager::finalize restatement.  Returns (rtype[27], rrho[27], ru[27,3])."""
    face_type = list(face_type) + [BC["Periodic"]] * (6 - len(face_type))
    fr = np.ones(6) if face_rho is None else np.asarray(face_rho, dtype=np.float64)
    fu = np.zeros((6, 3)) if face_u is None else np.ascontiguousarray(face_u, dtype=np.float64)
    rtype = (ctypes.c_int * 27)()
    rrho = np.zeros(27)
    ru = np.zeros((27, 3))
    rc = lib().orc_resolve_bc(dim, _i(face_type), _d(np.ascontiguousarray(fr)), _d(fu),
                              rtype, _d(rrho), _d(ru))
    if rc != 0:
        raise ValueError("conflicting boundary conditions at an edge/corner")
    return np.array(list(rtype), dtype=np.int64), rrho, ru


def step(lat, fin, fout, n, omega, F=None, nthreads=1):
    Fa, Fp = _force(F)
    lib().orc_step(lat.id, _d(fin), _d(fout), _i(n3(n)), omega, Fp, nthreads)


def step_masked(lat, fin, fout, n, omega, mask):
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    lib().orc_step_masked(lat.id, _d(fin), _d(fout), _i(n3(n)), omega, m.ctypes.data)


def run_masked(lat, f, n, omega, periodic, resolved, mask, nsteps):
    """advance() with solid obstacles: masked step, swap, ghost fill."""
    fa = np.ascontiguousarray(f, dtype=np.float64).copy()
    fb = np.full_like(fa, np.nan)
    for _ in range(nsteps):
        step_masked(lat, fa, fb, n, omega, mask)
        fa, fb = fb, fa
        fill_periodic(lat, fa, n, periodic)
        apply_bc(lat, fa, n, periodic, resolved)
    return fa


def fill_periodic(lat, f, n, periodic):
    lib().orc_fill_periodic(lat.id, _d(f), _i(n3(n)), _i(list(periodic) + [0] * (3 - len(periodic))))


def apply_bc(lat, f, n, periodic, resolved):
    rtype, rrho, ru = resolved
    rc = lib().orc_apply_bc(lat.id, _d(f), _i(n3(n)),
                            _i(list(periodic) + [0] * (3 - len(periodic))),
                            _i(rtype), _d(np.ascontiguousarray(rrho)),
                            _d(np.ascontiguousarray(ru)))
    if rc != 0:
        raise ValueError("unsupported boundary type")


def stream(lat, fin, fout, n):
    lib().orc_stream(lat.id, _d(fin), _d(fout), _i(n3(n)))


def moments(lat, f, n, F=None, streamed=True):
    nn = n3(n)
    rho = np.zeros((nn[2], nn[1], nn[0]))
    vel = np.zeros((lat.D, nn[2], nn[1], nn[0]))
    Fa, Fp = _force(F)
    lib().orc_moments(lat.id, _d(f), _i(nn), Fp, _d(rho), _d(vel), 1 if streamed else 0)
    return rho, vel


def init_equilibrium(lat, rho, vel, n, F=None):
    """setInitialConditions restatement.  rho (n2,n1,n0), vel (D,n2,n1,n0) are
    copied; returns (f, rho_out, vel_out)."""
    rho = np.ascontiguousarray(rho, dtype=np.float64).copy()
    vel = np.ascontiguousarray(vel, dtype=np.float64).copy()
    f = new_fab(lat, n)
    Fa, Fp = _force(F)
    lib().orc_init_equilibrium(lat.id, _d(rho), _d(vel), _d(f), _i(n3(n)), Fp)
    return f, rho, vel


def run(lat, f, n, omega, F, periodic, resolved, nsteps, nthreads=0):
    """advance() x nsteps starting from a ghost-filled fab; returns new fab."""
    rtype, rrho, ru = resolved
    fa = np.ascontiguousarray(f, dtype=np.float64).copy()
    fb = np.full_like(fa, np.nan)
    Fa, Fp = _force(F)
    r = lib().orc_run(lat.id, _d(fa), _d(fb), _i(n3(n)), omega, Fp,
                      _i(list(periodic) + [0] * (3 - len(periodic))), _i(rtype),
                      _d(np.ascontiguousarray(rrho)), _d(np.ascontiguousarray(ru)),
                      nsteps, nthreads)
    if r < 0:
        raise ValueError("oracle run failed")
    return fa if r == 0 else fb


def time_run_boxes(lat, f, n, mgs, omega, F, nwarm, nsteps, nthreads=0):
    """The fully periodic advance() loop on boxes of at most mgs cells per
    side (the reference's max_grid_size layout), OpenMP over boxes; f (a
    ghosted fab) is advanced in place.  Returns the seconds of the nsteps
    timed steps."""
    Fa, Fp = _force(F)
    L = lib()
    L.orc_time_run_boxes.restype = ctypes.c_double
    L.orc_time_run_boxes.argtypes = [ctypes.c_int, _dp, _ip, ctypes.c_int, ctypes.c_double, _dp,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int]
    assert f.flags.c_contiguous and f.dtype == np.float64
    t = L.orc_time_run_boxes(lat.id, _d(f), _i(n3(n)), mgs, omega, Fp, nwarm, nsteps, nthreads)
    if t < 0:
        raise ValueError("oracle boxed run failed")
    return t


def time_run(lat, fa, fb, n, omega, F, periodic, resolved, nsteps, nthreads=0):
    rtype, rrho, ru = resolved
    Fa, Fp = _force(F)
    return lib().orc_time_run(lat.id, _d(fa), _d(fb), _i(n3(n)), omega, Fp,
                              _i(list(periodic) + [0] * (3 - len(periodic))), _i(rtype),
                              _d(np.ascontiguousarray(rrho)), _d(np.ascontiguousarray(ru)),
                              nsteps, nthreads)
