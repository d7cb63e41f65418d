"""Python mirror of sl3dpp::detail's band solvers
(include/sl3dpp/detail/BandSolvers.hpp:24-49) over the C ABI, on device
arrays (torch tensors on the HIP device: plumbing for memory and streams).
Names and argument meaning follow the reference: band storage column-major,
pivots 0-based, right-hand sides n x nrhs column-major."""

import ctypes

from . import check, lib

_p = ctypes.c_void_p


def _ptr(t):
    return _p(t.data_ptr())


def band_cholesky_factor(ab, n, kd):
    """BandCholeskyFactor: factorises in place; returns LAPACK's info."""
    import torch

    torch.cuda.synchronize()
    info = ctypes.c_int(0)
    check(lib.sl3_band_cholesky_factor(_ptr(ab), n, kd, ctypes.byref(info)))
    return info.value


def band_cholesky_solve(ab, n, kd, b, nrhs):
    """BandCholeskySolve: b (n x nrhs column-major) becomes X."""
    import torch

    torch.cuda.synchronize()
    check(lib.sl3_band_cholesky_solve(_ptr(ab), n, kd, _ptr(b), nrhs))


def band_lu_factor(ab, n, kl, ku, pivots):
    """BandLuFactor: factorises in place, 0-based pivots; returns info."""
    import torch

    torch.cuda.synchronize()
    info = ctypes.c_int(0)
    check(lib.sl3_band_lu_factor(_ptr(ab), n, kl, ku, _ptr(pivots),
                                 ctypes.byref(info)))
    return info.value


def band_lu_solve(ab, n, kl, ku, pivots, b, nrhs):
    """BandLuSolve: b (n x nrhs column-major) becomes X."""
    import torch

    torch.cuda.synchronize()
    check(lib.sl3_band_lu_solve(_ptr(ab), n, kl, ku, _ptr(pivots), _ptr(b), nrhs))
