"""ctypes binding of the step engine's C ABI (include/mfa_lbm.h).

This is the Python-side equivalent of the binding a C++ or Fortran caller
would write: plain structs and pointers, no torch types in the ABI.  torch is
used only to own device memory and to name the HIP stream.  Loading fails
loudly when the built library is missing; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmfa_lbm.so")

LATTICE = {"D2Q9": 0, "D3Q15": 1, "D3Q19": 2, "D3Q27": 3}
LATTICE_DQ = {0: (2, 9), 1: (3, 15), 2: (3, 19), 3: (3, 27)}
BC = {"MovingWall": 0, "NoSlip": 1, "Inflow": 2, "Outflow": 3, "PressureOutlet": 4,
      "Periodic": 5}
DTYPE = {"f64": 0, "f32": 1, "f16dev": 2, "bf16dev": 3}
RED = {"mass": 0, "momentum": 1, "comp_sum": 2, "health": 3}

EXPORTS = ["mfa_version", "mfa_last_error", "mfa_resolve_bc", "mfa_step", "mfa_step2", "mfa_pull_collide",
           "mfa_fill_periodic", "mfa_apply_bc", "mfa_moments", "mfa_init_equilibrium",
           "mfa_reduce", "mfa_residual", "mfa_set_launch", "mfa_get_launch", "mfa_set_launch_dev",
           "mfa_get_launch_dev", "mfa_stream_copy",
           "mfa_convert", "mfa_level_create", "mfa_level_destroy", "mfa_level_bind",
           "mfa_level_fill_boundary", "mfa_level_apply_bc", "mfa_level_ghost_cells",
           "mfa_level_step", "mfa_reduce_noalloc", "mfa_level_step_direct",
           "mfa_level_direct_available", "mfa_level_set_direct_launch"]


class MfaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mfa status {code}: {msg}")
        self.code = code


class Field(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("n", ctypes.c_int * 3), ("ngrow", ctypes.c_int),
                ("sy", ctypes.c_longlong), ("sz", ctypes.c_longlong), ("sq", ctypes.c_longlong),
                ("ncomp", ctypes.c_int), ("dtype", ctypes.c_int)]


class Launch(ctypes.Structure):
    _fields_ = [("block_x", ctypes.c_int), ("block_y", ctypes.c_int),
                ("cells_per_thread", ctypes.c_int), ("nt_loads", ctypes.c_int)]


class GeomBC(ctypes.Structure):
    _fields_ = [("lattice", ctypes.c_int), ("periodic", ctypes.c_int * 3),
                ("resolved_type", ctypes.c_int * 27), ("resolved_rho", ctypes.c_double * 27),
                ("resolved_u", (ctypes.c_double * 3) * 27)]


def load_library(path):
    """dlopen one of this package's libraries.

    torch ships its own HIP runtime under the same soname (libamdhip64.so.7)
    as /opt/rocm's, and the dynamic loader binds every later user of that
    soname to whichever copy was loaded first.  torch is imported before the
    library so that the process has the one runtime torch was built against;
    loaded the other way round, torch later sees no device.
    """
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    return ctypes.CDLL(path)


_lib = None


def lib():
    """Load libmfa_lbm.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m muphfasa_amd.build` "
                "(there is no CPU fallback for the step engine)")
        L = load_library(LIB_PATH)
        F = ctypes.POINTER(Field)
        G = ctypes.POINTER(GeomBC)
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        L.mfa_version.restype = ctypes.c_char_p
        L.mfa_last_error.restype = ctypes.c_char_p
        L.mfa_resolve_bc.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), dp, dp, G]
        L.mfa_step.argtypes = [F, F, G, ctypes.c_double, dp, vp]
        L.mfa_step2.argtypes = [F, F, G, ctypes.c_double, dp, vp]
        L.mfa_pull_collide.argtypes = [F, F, ctypes.c_int, ctypes.c_double, dp, vp]
        L.mfa_fill_periodic.argtypes = [F, G, vp]
        L.mfa_apply_bc.argtypes = [F, G, vp]
        L.mfa_moments.argtypes = [F, F, F, G, dp, ctypes.c_int, vp]
        L.mfa_init_equilibrium.argtypes = [F, F, F, ctypes.c_int, dp, vp]
        L.mfa_reduce.argtypes = [F, ctypes.c_int, ctypes.c_int, vp, vp]
        L.mfa_reduce_noalloc.argtypes = [F, ctypes.c_int, ctypes.c_int, vp, vp]
        L.mfa_residual.argtypes = [F, F, vp, vp]
        L.mfa_set_launch.argtypes = [ctypes.c_int, ctypes.POINTER(Launch)]
        L.mfa_get_launch.argtypes = [ctypes.c_int, ctypes.POINTER(Launch)]
        L.mfa_set_launch_dev.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(Launch)]
        L.mfa_get_launch_dev.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(Launch)]
        L.mfa_stream_copy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int, vp]
        L.mfa_convert.argtypes = [F, F, ctypes.c_int, vp]
        ip = ctypes.POINTER(ctypes.c_int)
        L.mfa_level_create.argtypes = [ctypes.c_int, ip, ip, ctypes.c_int, ip, ip, ctypes.POINTER(vp)]
        L.mfa_level_destroy.argtypes = [vp]
        L.mfa_level_bind.argtypes = [vp, ctypes.c_int, F]
        L.mfa_level_fill_boundary.argtypes = [vp, ctypes.c_int, vp]
        L.mfa_level_apply_bc.argtypes = [vp, ctypes.c_int, G, vp]
        L.mfa_level_ghost_cells.argtypes = [vp]
        L.mfa_level_step.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp, vp]
        L.mfa_level_step_direct.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp, vp]
        L.mfa_level_direct_available.argtypes = [vp]
        L.mfa_level_set_direct_launch.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        for name in EXPORTS[2:]:
            getattr(L, name).restype = ctypes.c_int
        L.mfa_level_ghost_cells.restype = ctypes.c_longlong
        _lib = L
    return _lib


def check(rc: int):
    if rc != 0:
        raise MfaError(rc, lib().mfa_last_error().decode())


def _dbl(vals, n=3):
    if vals is None:
        return None
    a = (ctypes.c_double * n)()
    for i, v in enumerate(list(vals)[:n]):
        a[i] = float(v)
    return a


def resolve_bc(lattice: int, face_type, face_rho=None, face_u=None) -> GeomBC:
    """BCManager::finalize on the host (C ABI mfa_resolve_bc)."""
    D = LATTICE_DQ[lattice][0]
    ft = list(face_type) + [BC["Periodic"]] * (6 - len(face_type))
    fta = (ctypes.c_int * 6)(*ft)
    fr = _dbl(face_rho if face_rho is not None else [1.0] * 6, 6)
    fu = (ctypes.c_double * 18)()
    if face_u is not None:
        for f in range(6):
            for d in range(3):
                fu[3 * f + d] = float(face_u[f][d])
    g = GeomBC()
    check(lib().mfa_resolve_bc(lattice, fta, fr, fu, ctypes.byref(g)))
    return g


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def field(t, n, ngrow: int, lattice_dim: int, dtype: int = 0, layout: str = "component") -> Field:
    """mfa_field over a contiguous tensor whose extents are n_d + 2*ngrow in
    the lattice's dimensions (z has no ghost in 2-D) or larger (padding).

    layout "component": shape (ncomp, e2, e1, e0), the MultiFab's order (one
    plane per component).  layout "row": shape (e2, e1, ncomp, e0), the Q
    rows of one (y, z) side by side (sq = e0, sy = ncomp*e0)."""
    assert t.is_contiguous()
    if layout == "component":
        ncomp, e2, e1, e0 = t.shape
        sy, sz, sq = e0, e0 * e1, e0 * e1 * e2
    elif layout == "row":
        e2, e1, ncomp, e0 = t.shape
        sy, sz, sq = ncomp * e0, ncomp * e0 * e1, e0
    else:
        raise ValueError(f"unknown layout {layout!r}")
    n = list(n) + [1] * (3 - len(n))
    gz = ngrow if lattice_dim == 3 else 0
    assert e0 >= n[0] + 2 * ngrow and e1 >= n[1] + 2 * ngrow and e2 >= n[2] + 2 * gz
    f = Field()
    esz = t.element_size()
    off = ngrow + ngrow * sy + gz * sz
    f.data = t.data_ptr() + off * esz
    for d in range(3):
        f.n[d] = n[d]
    f.ngrow = ngrow
    f.sy, f.sz, f.sq = sy, sz, sq
    f.ncomp = ncomp
    f.dtype = dtype
    return f


def row_shape(Q: int, n, ngrow: int, lattice_dim: int):
    """Tensor shape of a row-interleaved population field (for field(..., layout="row"))."""
    n = list(n) + [1] * (3 - len(n))
    gz = ngrow if lattice_dim == 3 else 0
    return (n[2] + 2 * gz, n[1] + 2 * ngrow, Q, n[0] + 2 * ngrow)


def step(fin: Field, fout: Field, g: GeomBC, omega: float, force=None, stream=None):
    check(lib().mfa_step(ctypes.byref(fin), ctypes.byref(fout), ctypes.byref(g), omega,
                         _dbl(force), stream_ptr(stream)))


def step2(fin: Field, fout: Field, g: GeomBC, omega: float, force=None, stream=None):
    """Two steps fin -> fout in one pass, bit-identical to two step() calls
    (FP64: every lattice, periodic or with walls; FP32 storage: 2-D boxes and
    fully periodic 3-D boxes; MfaError with code -2 elsewhere)."""
    check(lib().mfa_step2(ctypes.byref(fin), ctypes.byref(fout), ctypes.byref(g), omega,
                          _dbl(force), stream_ptr(stream)))


def pull_collide(fin: Field, fout: Field, lattice: int, omega: float, force=None, stream=None):
    check(lib().mfa_pull_collide(ctypes.byref(fin), ctypes.byref(fout), lattice, omega,
                                 _dbl(force), stream_ptr(stream)))


def fill_periodic(f: Field, g: GeomBC, stream=None):
    check(lib().mfa_fill_periodic(ctypes.byref(f), ctypes.byref(g), stream_ptr(stream)))


def apply_bc(f: Field, g: GeomBC, stream=None):
    check(lib().mfa_apply_bc(ctypes.byref(f), ctypes.byref(g), stream_ptr(stream)))


def moments(f: Field, rho: Field, vel: Field, g: GeomBC, force=None, streamed=True,
            stream=None):
    """streamed: True (pull with inline wrap and walls), "halo" (pull through
    f's filled ghost cells) or False (unstreamed)."""
    mode = 2 if streamed == "halo" else (1 if streamed else 0)
    check(lib().mfa_moments(ctypes.byref(f), ctypes.byref(rho), ctypes.byref(vel),
                            ctypes.byref(g), _dbl(force), mode, stream_ptr(stream)))


class Level:
    """Boxes of one lattice on one device (include/mfa_level.h): the ghost
    exchange between them and the physical boundary fill of each box."""

    def __init__(self, lattice: int, domain_n, periodic, box_lo, box_n):
        nb = len(box_lo)
        lo = (ctypes.c_int * (3 * nb))(*[int(v) for b in box_lo for v in (list(b) + [0, 0])[:3]])
        nn = (ctypes.c_int * (3 * nb))(*[int(v) for b in box_n for v in (list(b) + [1, 1])[:3]])
        dn = (ctypes.c_int * 3)(*(list(domain_n) + [1, 1])[:3])
        per = (ctypes.c_int * 3)(*(list(periodic) + [1, 1])[:3])
        h = ctypes.c_void_p()
        check(lib().mfa_level_create(lattice, dn, per, nb, lo, nn, ctypes.byref(h)))
        self.h = h
        self.nboxes = nb

    def close(self):
        if self.h:
            lib().mfa_level_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ghost_cells(self) -> int:
        return int(lib().mfa_level_ghost_cells(self.h))

    def bind(self, slot: int, fields):
        arr = (Field * len(fields))(*fields)
        check(lib().mfa_level_bind(self.h, slot, arr))

    def fill_boundary(self, slot: int, stream=None):
        check(lib().mfa_level_fill_boundary(self.h, slot, stream_ptr(stream)))

    def apply_bc(self, slot: int, g: GeomBC, stream=None):
        check(lib().mfa_level_apply_bc(self.h, slot, ctypes.byref(g), stream_ptr(stream)))

    def step(self, slot_in: int, slot_out: int, omega: float, force=None, stream=None):
        check(lib().mfa_level_step(self.h, slot_in, slot_out, omega, _dbl(force), stream_ptr(stream)))

    @property
    def direct_available(self) -> bool:
        """Whether step_direct serves this layout (periodic domain, boxes on a
        tensor-product grid)."""
        return bool(lib().mfa_level_direct_available(self.h))

    def set_direct_launch(self, cells_per_thread: int, nontemporal: int):
        check(lib().mfa_level_set_direct_launch(self.h, cells_per_thread, nontemporal))

    def step_direct(self, slot_in: int, slot_out: int, omega: float, force=None, stream=None):
        """The step reading other boxes' valid cells directly (no exchange)."""
        check(lib().mfa_level_step_direct(self.h, slot_in, slot_out, omega, _dbl(force),
                                          stream_ptr(stream)))


def init_equilibrium(rho: Field, vel: Field, f: Field, lattice: int, force=None, stream=None):
    check(lib().mfa_init_equilibrium(ctypes.byref(rho), ctypes.byref(vel), ctypes.byref(f),
                                     lattice, _dbl(force), stream_ptr(stream)))


def reduce(f: Field, lattice: int, what: int, out_ptr: int, stream=None):
    check(lib().mfa_reduce(ctypes.byref(f), lattice, what, ctypes.c_void_p(out_ptr),
                           stream_ptr(stream)))


def reduce_noalloc(f: Field, lattice: int, what: int, out_ptr: int, stream=None):
    """reduce() with its partial sums in the library's file-scope device
    array instead of an allocated workspace (mfa_reduce_noalloc)."""
    check(lib().mfa_reduce_noalloc(ctypes.byref(f), lattice, what, ctypes.c_void_p(out_ptr),
                                   stream_ptr(stream)))


def residual(a: Field, b: Field, out_ptr: int, stream=None):
    check(lib().mfa_residual(ctypes.byref(a), ctypes.byref(b), ctypes.c_void_p(out_ptr),
                             stream_ptr(stream)))


def set_launch(lattice: int, block_x=None, block_y=1, cells_per_thread=1, nt_loads=0):
    """Set mfa_step's launch configuration for a lattice (block_x None resets)."""
    if block_x is None:
        check(lib().mfa_set_launch(lattice, None))
        return
    c = Launch(block_x, block_y, cells_per_thread, nt_loads)
    check(lib().mfa_set_launch(lattice, ctypes.byref(c)))


def get_launch(lattice: int):
    c = Launch()
    check(lib().mfa_get_launch(lattice, ctypes.byref(c)))
    return (c.block_x, c.block_y, c.cells_per_thread, c.nt_loads)


def set_launch_dev(lattice: int, dtype: int, block_x=None, block_y=1, cells_per_thread=2, nt_loads=1):
    """Launch configuration of the reduced-storage step (block_x None resets);
    cells_per_thread 2 = two x-adjacent cells per thread with paired accesses."""
    if block_x is None:
        check(lib().mfa_set_launch_dev(lattice, dtype, None))
        return
    c = Launch(block_x, block_y, cells_per_thread, nt_loads)
    check(lib().mfa_set_launch_dev(lattice, dtype, ctypes.byref(c)))


def get_launch_dev(lattice: int, dtype: int):
    c = Launch()
    check(lib().mfa_get_launch_dev(lattice, dtype, ctypes.byref(c)))
    return (c.block_x, c.block_y, c.cells_per_thread, c.nt_loads)


def stream_copy(src_ptr: int, dst_ptr: int, nbytes: int, nontemporal=True, stream=None):
    check(lib().mfa_stream_copy(ctypes.c_void_p(src_ptr), ctypes.c_void_p(dst_ptr), nbytes,
                                1 if nontemporal else 0, stream_ptr(stream)))


def convert(src: Field, dst: Field, lattice: int, stream=None):
    check(lib().mfa_convert(ctypes.byref(src), ctypes.byref(dst), lattice, stream_ptr(stream)))
