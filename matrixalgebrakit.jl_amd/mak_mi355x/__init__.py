"""MI355X-native QR / polar backend for MatrixAlgebraKit.jl (host-side mirror).

The compute path is libmak_mi355x.so (hand-written HIP kernels for gfx950,
C ABI in include/mak_mi355x.h); this package only checks arguments, allocates
outputs / workspace with torch, and calls the C ABI on the current stream.
"""
from .api import *  # noqa: F401,F403
from .api import __all__ as _api_all
from ._lib import LIB_PATH, LibraryNotBuilt, declared_symbols, load

__all__ = list(_api_all) + ["LIB_PATH", "LibraryNotBuilt", "declared_symbols", "load"]
from .pullbacks import *  # noqa: F401,F403,E402
from .pullbacks import __all__ as _pb_all  # noqa: E402

__all__ += list(_pb_all)
from .batched import *  # noqa: F401,F403,E402
from .batched import __all__ as _b_all  # noqa: E402

__all__ += list(_b_all)
