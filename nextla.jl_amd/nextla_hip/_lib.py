"""ctypes binding of libnextla_hip.so (the C ABI declared in include/nextla_hip.h).

The library is loaded from nextla.jl_amd/lib/.  There is no fallback: if the shared
object is missing or cannot be loaded, every call raises.
"""
import ctypes
import os
import re

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "lib", "libnextla_hip.so")
HEADER = os.path.join(os.path.dirname(_PKG), "include", "nextla_hip.h")

i64 = ctypes.c_int64
ch = ctypes.c_char
vp = ctypes.c_void_p
dbl = ctypes.c_double
flt = ctypes.c_float

# argument kinds per routine (type-generic; 'R' = the real scalar type)
_SIGS = {
    "larfg": [i64, vp, vp, i64, vp, vp],
    "geqr2": [i64, i64, vp, i64, vp, vp, vp],
    "larft": [ch, ch, i64, i64, vp, i64, vp, vp, i64, vp],
    "geqrt": [i64, i64, i64, vp, i64, vp, i64, vp, vp, vp],
    "larfb": [ch, ch, ch, ch, i64, i64, i64, vp, i64, vp, i64, vp, i64, vp, i64, vp],
    "unmqr": [ch, ch, i64, i64, i64, i64, vp, i64, i64, vp, i64, vp, i64, vp, vp],
    "tsqrt": [i64, i64, i64, vp, i64, i64, vp, i64, vp, i64, vp, vp, vp],
    "tsmqr": [ch, ch, i64, i64, i64, i64, i64, i64, vp, i64, vp, i64, vp, i64, i64, vp, i64, i64, vp, vp],
    "parfb": [ch, ch, ch, ch, i64, i64, i64, i64, i64, i64, vp, i64, vp, i64, vp, i64, vp, i64, vp, i64, vp],
    "ttqrt": [i64, i64, i64, vp, i64, vp, i64, vp, i64, vp, vp, vp],
    "ttmqr": [ch, ch, i64, i64, i64, i64, i64, i64, vp, i64, vp, i64, vp, i64, vp, i64, vp, vp],
    "pamm": [ch, ch, ch, ch, i64, i64, i64, i64, vp, i64, vp, i64, vp, i64, vp, i64, vp],
    "pemv": [ch, ch, i64, i64, i64, "R", vp, i64, vp, "R", vp, vp, vp],
    "rectrxm": [ch, ch, ch, "R", ch, i64, i64, vp, i64, vp, i64, i64, vp],
    "trsm_base": [ch, ch, i64, i64, vp, i64, vp, i64, vp],
    "trmm_base": [ch, ch, i64, i64, vp, i64, vp, i64, vp],
    "gemm_acc": [i64, i64, i64, "R", vp, i64, vp, i64, vp, i64, vp],
    "tile_geqrf": [i64, i64, i64, i64, vp, i64, vp, i64, vp],
    "tile_geqrf_calls": [i64, i64, i64, i64, vp, i64, vp, i64, vp],
    "tile_ormqr": [ch, i64, i64, i64, i64, i64, vp, i64, vp, i64, vp, i64, vp],
    "trsm_rrule": [ch, ch, ch, "R", i64, i64, vp, i64, vp, i64, vp, i64, vp, i64, i64, vp],
    "trmm_rrule": [ch, ch, ch, "R", i64, i64, vp, i64, vp, i64, vp, i64, vp, i64, i64, vp],
    "qr_rrule": [i64, i64, i64, i64, vp, i64, vp, i64, vp, i64, vp, i64, vp],
    "tile_gels": [i64, i64, i64, i64, vp, i64, vp, vp, ch, i64, dbl, vp, vp, vp, vp],
    "batched_qr_trsm": [i64, vp, vp, ctypes.c_int, vp],
    "batched_autotune": [i64, vp, vp, i64, vp, i64, vp, vp, vp],
    "getrf2": [i64, i64, vp, i64, vp, vp, vp],
    "laswp": [i64, vp, i64, i64, i64, vp, i64, vp],
}
# Float64-only entry points
_SIGS_D = {
    "tile_geqrf_mixed": [i64, i64, i64, i64, vp, i64, vp, i64, ch, vp],
    "rectrsm_mixed": [ch, ch, ch, dbl, i64, i64, vp, i64, vp, i64, i64, i64, dbl, ch, vp, vp, vp],
    "rectrxm_noalloc": [ch, ch, ch, dbl, ch, i64, i64, vp, i64, vp, i64, i64, vp],
}

# the QR-family routines also exist for complex element types ('z' ComplexF64, 'c' ComplexF32)
QR_FAMILY = ("larfg", "geqr2", "larft", "geqrt", "larfb", "unmqr", "tsqrt", "tsmqr", "parfb", "ttqrt", "ttmqr",
             "pamm", "pemv", "tile_geqrf_calls")


class ZComplex(ctypes.Structure):
    """nla_zcomplex: a ComplexF64 passed by value."""
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


class CComplex(ctypes.Structure):
    """nla_ccomplex: a ComplexF32 passed by value."""
    _fields_ = [("re", ctypes.c_float), ("im", ctypes.c_float)]


_lib = None


class ArgumentError(ValueError):
    """Raised where NextLA.jl throws ArgumentError (negative status from the library)."""


class HipError(RuntimeError):
    """A HIP runtime error reported by the library (positive status)."""


def load():
    """Load libnextla_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libnextla_hip.so is not built ({LIB_PATH}); run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    lib.nla_last_error.restype = ctypes.c_char_p
    lib.nla_version.restype = ctypes.c_int
    lib.nla_synchronize.argtypes = [vp]
    lib.nla_scratch_bytes.restype = i64
    lib.nla_prof_enable.argtypes = [ctypes.c_int]
    lib.nla_prof_reset.restype = None
    lib.nla_prof_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(dbl), ctypes.POINTER(i64)]
    lib.nla_tile_qr_schedule.argtypes = [i64, i64, i64, ctypes.POINTER(ctypes.c_int32), i64]
    lib.nla_tile_qr_schedule.restype = i64
    for name, sig in _SIGS.items():
        kinds = (("d", dbl), ("s", flt)) + ((("z", ZComplex), ("c", CComplex)) if name in QR_FAMILY else ())
        for p, real in kinds:
            f = getattr(lib, f"nla_{p}{name}")
            f.argtypes = [real if a == "R" else a for a in sig]
            f.restype = ctypes.c_int
    for p, real in (("d", dbl), ("s", flt)):
        f = getattr(lib, f"nla_h{p}rectrxm")
        f.argtypes = [ch, ch, ch, real, ch, i64, i64, vp, i64, vp, i64, i64, vp]
        f.restype = ctypes.c_int
    for name, sig in _SIGS_D.items():
        f = getattr(lib, f"nla_d{name}")
        f.argtypes = sig
        f.restype = ctypes.c_int
    _lib = lib
    return lib


def header_symbols():
    """Every function name declared in include/nextla_hip.h."""
    with open(HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(nla_[a-z0-9_]+)\s*\(", text)))


def check(status, fn):
    if status == 0:
        return
    msg = load().nla_last_error().decode(errors="replace")
    if status < 0:
        raise ArgumentError(msg or f"{fn}: illegal argument {-status}")
    raise HipError(f"{fn}: {msg} (status {status})")


def call(prefix, name, *args):
    lib = load()
    fn = f"nla_{prefix}{name}"
    st = getattr(lib, fn)(*args)
    check(st, fn)
    return st


def prof_enable(on=True):
    load().nla_prof_enable(1 if on else 0)


def prof_reset():
    load().nla_prof_reset()


def prof_read(kernel):
    """(total milliseconds, launches) recorded for one kernel family while profiling was on."""
    ms = dbl(0)
    n = i64(0)
    check(load().nla_prof_read(kernel.encode(), ctypes.byref(ms), ctypes.byref(n)), "nla_prof_read")
    return ms.value, n.value


def tile_qr_schedule(m, n, nb):
    """The claim order of the tile-QR persistent launch for an m x n matrix: a list of
    ("P", k, stage) and ("W", k, strip_group) tuples (host-only, no GPU call)."""
    lib = load()
    cnt = lib.nla_tile_qr_schedule(m, n, nb, None, 0)
    if cnt < 0:
        raise ArgumentError(lib.nla_last_error().decode())
    buf = (ctypes.c_int32 * cnt)()
    lib.nla_tile_qr_schedule(m, n, nb, buf, cnt)
    return [("P" if (c >> 30) == 0 else "W", (c >> 18) & 0xFFF, c & 0x3FFFF) for c in buf]
