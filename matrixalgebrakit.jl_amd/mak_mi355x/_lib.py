"""ctypes binding of libmak_mi355x.so (the C ABI declared in include/mak_mi355x.h).

The library is REQUIRED: importing the compute entry points without it raises
immediately; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libmak_mi355x.so"
HEADER = Path(__file__).resolve().parents[2] / "include" / "mak_mi355x.h"

_vp = C.c_void_p
_i = C.c_int
_sz = C.c_size_t
_d = C.c_double
_ll = C.c_longlong

# name -> (restype, argtypes); kept in the order of include/mak_mi355x.h
SIGNATURES = {
    "mak_qr_workspace_size": (_i, [_i, _i, _i, _i, C.POINTER(_sz)]),
    "mak_qr_compact_d": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _i, _vp, _sz]),
    "mak_qr_compact_z": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _i, _vp, _sz]),
    "mak_qr_full_d": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _i, _vp, _sz]),
    "mak_qr_full_z": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _i, _vp, _sz]),
    "mak_qr_compact_wy_d": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _sz]),
    "mak_qr_compact_wy_z": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _sz]),
    "mak_qr_compact_noalloc_d": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _i]),
    "mak_lq_workspace_size": (_i, [_i, _i, _i, _i, C.POINTER(_sz)]),
    "mak_lq_compact_d": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _i, _vp, _sz]),
    "mak_lq_compact_z": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _i, _vp, _sz]),
    "mak_lq_full_d": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _i, _vp, _sz]),
    "mak_lq_full_z": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _i, _vp, _sz]),
    "mak_null_workspace_size": (_i, [_i, _i, _i, _i, C.POINTER(_sz)]),
    "mak_qr_null_d": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _sz]),
    "mak_qr_null_z": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _sz]),
    "mak_lq_null_d": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _sz]),
    "mak_lq_null_z": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _sz]),
    "mak_polar_workspace_size": (_i, [_i, _i, _i, C.POINTER(_sz)]),
    "mak_left_polar_d": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _i, _i, _d, _vp, _sz, _vp]),
    "mak_left_polar_z": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _i, _i, _d, _vp, _sz, _vp]),
    "mak_right_polar_d": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _i, _i, _d, _vp, _sz, _vp]),
    "mak_right_polar_z": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _i, _i, _d, _vp, _sz, _vp]),
    "mak_svd_polar_workspace_size": (_i, [_i, _i, _i, C.POINTER(_sz)]),
    "mak_svd_polar_d": (_i, [_vp, _i, _i, _vp, _i, _vp, _vp, _i, _vp, _i, _i, _vp, _sz, _vp]),
    "mak_svd_polar_z": (_i, [_vp, _i, _i, _vp, _i, _vp, _vp, _i, _vp, _i, _i, _vp, _sz, _vp]),
    "mak_polar_pushforward_workspace_size": (_i, [_i, _i, _i, _i, C.POINTER(_sz)]),
    "mak_left_polar_pushforward_d": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _sz]),
    "mak_left_polar_pushforward_z": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _sz]),
    "mak_right_polar_pushforward_d": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _sz]),
    "mak_right_polar_pushforward_z": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _sz]),
    "mak_pullback_workspace_size": (_i, [_i, _i, _i, _i, _i, C.POINTER(_sz)]),
    "mak_qr_pullback_d": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _i, _ll, _vp, _ll, _vp, _i, _ll, _vp, _i, _ll,
                               _vp, _i, _ll, _vp, _i, _ll, _vp, _sz, _vp]),
    "mak_qr_pullback_z": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _i, _ll, _vp, _ll, _vp, _i, _ll, _vp, _i, _ll,
                               _vp, _i, _ll, _vp, _i, _ll, _vp, _sz, _vp]),
    "mak_qr_full_cotangent_d": (_i, [_vp, _i, _i, _i, _i, _vp, _i, _ll, _vp, _i, _ll, _vp, _i, _ll, _vp, _vp,
                                     _sz]),
    "mak_qr_full_cotangent_z": (_i, [_vp, _i, _i, _i, _i, _vp, _i, _ll, _vp, _i, _ll, _vp, _i, _ll, _vp, _vp,
                                     _sz]),
    "mak_left_polar_pullback_d": (_i, [_vp, _i, _i, _i, _vp, _i, _ll, _vp, _i, _ll, _vp, _i, _ll, _vp, _i, _ll,
                                       _vp, _i, _ll, _i, _d, _vp, _sz, _vp]),
    "mak_left_polar_pullback_z": (_i, [_vp, _i, _i, _i, _vp, _i, _ll, _vp, _i, _ll, _vp, _i, _ll, _vp, _i, _ll,
                                       _vp, _i, _ll, _i, _d, _vp, _sz, _vp]),
    "mak_qr_batched_workspace_size": (_i, [_i, _i, _vp, _vp, _i, C.POINTER(_sz)]),
    "mak_qr_batched_plan_create": (_i, [C.POINTER(_vp), _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i,
                                        _vp, _sz]),
    "mak_qr_batched_execute": (_i, [_vp, _vp]),
    "mak_qr_batched_plan_info": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "mak_qr_batched_plan_destroy": (_i, [_vp]),
    "mak_qr_compact_batched": (_i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _sz]),
}


class PolarInfo(C.Structure):
    """mak_polar_info (include/mak_mi355x.h)."""
    _fields_ = [("iterations", C.c_int), ("qr_steps", C.c_int), ("chol_steps", C.c_int),
                ("ns_steps", C.c_int), ("alpha", C.c_double), ("l0", C.c_double),
                ("l_final", C.c_double), ("last_diff", C.c_double), ("completed", C.c_int),
                ("poly_steps", C.c_int)]

    def asdict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LibraryNotBuilt(RuntimeError):
    pass


_LIB = None


def load(path: os.PathLike | None = None):
    """Load (once) and type the shared library; raise if it is missing."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise LibraryNotBuilt(
            f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(the HIP library is required; there is no CPU fallback)")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def declared_symbols(header: os.PathLike = HEADER):
    """Function names declared in include/mak_mi355x.h."""
    text = Path(header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"\bint\s+(mak_\w+)\s*\(", text)
