"""sl3dpp_amd: the MI355X (gfx950) implementation of the data-parallel core of
sl3dpp's 3D lateral-viscosity step, loaded over its C ABI
(include/sl3dpp_amd.h) with ctypes.

The HIP library is required: importing raises if libsl3dpp_amd.so is not
built, and there is no CPU fallback on any product path.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "libsl3dpp_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "sl3dpp_amd.h")

# One HIP runtime per process: torch bundles its own libamdhip64 and HSA
# runtime, and a process that loads this library first would then hold two
# (the library's from /opt/rocm and torch's), which do not share devices.
# Loading torch first lets the library bind to the runtime already loaded.
try:
    import torch  # noqa: F401
except ImportError:  # no torch: the library's own runtime from /opt/rocm
    torch = None

if not os.path.exists(LIBRARY_PATH):
    raise ImportError(
        f"{LIBRARY_PATH} is not built: run __graft_entry__.build() (hipcc, gfx950)"
    )

lib = ctypes.CDLL(LIBRARY_PATH)
lib.sl3_last_error.restype = ctypes.c_char_p
lib.sl3_version.restype = ctypes.c_char_p


class Sl3Error(RuntimeError):
    """A failed C-ABI call: status code and the library's message."""

    def __init__(self, status, message):
        super().__init__(f"sl3 status {status}: {message}")
        self.status = status


class InvalidArgument(Sl3Error, ValueError):
    """The reference's std::invalid_argument."""


class FactorisationError(Sl3Error):
    """The reference's std::runtime_error from a failed factorisation."""


def check(status):
    """Raises the exception the reference would throw for a status code."""
    if status == 0:
        return
    message = lib.sl3_last_error().decode(errors="replace")
    if status == -1:
        raise InvalidArgument(status, message)
    if status > 0:
        raise FactorisationError(status, message)
    raise Sl3Error(status, message)


def synchronize():
    check(lib.sl3_synchronize())


def version():
    return lib.sl3_version().decode()


from . import band  # noqa: E402,F401  (the band solvers' Python mirror)


def profile_enable(on=True):
    """Device-event timing of the library's kernels (off by default)."""
    check(lib.sl3_profile_enable(1 if on else 0))


def profile_select(names=None):
    """Times only the named kernels (an iterable of names; None: all)."""
    check(lib.sl3_profile_select(",".join(names).encode() if names else None))


def tuning_set(analysis_rows=0, synthesis_theta=0, band_panel_rows=0):
    """Fixes the Legendre kernels' tile shapes of operators made from now
    on (degrees per analysis workgroup 16, 32, 64 or 128; colatitudes per
    synthesis workgroup 32, 64, 112 or 208) and the 1D evolution's band
    panel (4, 8 or 16 rows); 0 lets the tuner choose."""
    check(lib.sl3_tuning_set(int(analysis_rows), int(synthesis_theta),
                             int(band_panel_rows)))


def profile_reset():
    check(lib.sl3_profile_reset())


def profile_read():
    """{kernel: (total_ms, launches)} since the last reset."""
    buf = ctypes.create_string_buffer(1 << 14)
    check(lib.sl3_profile_names(buf, len(buf)))
    out = {}
    for name in buf.value.decode().split():
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        check(lib.sl3_profile_read(name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        out[name] = (ms.value, n.value)
    return out
