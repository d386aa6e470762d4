"""MI355X (gfx950) lattice-Boltzmann step engine for MuPhFASa.

The compute path is libmfa_lbm.so (hand-written HIP behind the C ABI in
include/mfa_lbm.h); ``muphfasa_amd.abi`` binds it.  Nothing here falls back to
CPU arithmetic: if the library is missing, ``abi.lib()`` raises.
"""
__all__ = ["abi"]
