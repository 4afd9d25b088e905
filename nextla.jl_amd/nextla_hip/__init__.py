"""nextla_hip: MI355X (gfx950) implementation of NextLA.jl's tile-QR kernels and
recursive TRSM/TRMM behind NextLA's interface (see api.py)."""
from ._lib import ArgumentError, HipError, LIB_PATH, header_symbols, load, prof_enable, prof_read, prof_reset, tile_qr_schedule  # noqa: F401
from .api import *  # noqa: F401,F403
from . import autodiff, batched  # noqa: F401
