"""The C-ABI library loads and exports every function include/sl3dpp_amd.h
declares (no compute call: this runs without a GPU)."""

import ctypes
import os
import re

from conftest import ROOT


def declared_functions():
    text = open(os.path.join(ROOT, "include", "sl3dpp_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sl3_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    import sl3dpp_amd

    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(sl3dpp_amd.lib, n)]
    assert not missing, missing


def test_version_and_error_string_without_a_device():
    import sl3dpp_amd

    assert "gfx950" in sl3dpp_amd.version()
    assert isinstance(sl3dpp_amd.lib.sl3_last_error(), bytes)


def test_learned_entry_points_reject_bad_arguments_without_a_device():
    """Argument checks come before any device call: status -1 (the
    reference's std::invalid_argument) for an empty network, null handles
    and a negative warm-start guard's missing operator."""
    import sl3dpp_amd

    lib = sl3dpp_amd.lib
    out = ctypes.c_void_p()
    widths = (ctypes.c_int * 2)(4, 0)
    assert lib.sl3_mlp_create(1, widths, 16, ctypes.c_uint64(1), ctypes.byref(out)) == -1
    assert b"positive" in lib.sl3_last_error()
    assert lib.sl3_mlp_create(0, widths, 16, ctypes.c_uint64(1), ctypes.byref(out)) == -1
    assert lib.sl3_mlp_forward(None, 1, None, None) == -1
    assert lib.sl3_mlp_adam(None, ctypes.c_double(1e-3), ctypes.c_double(0.9),
                            ctypes.c_double(0.999), ctypes.c_double(1e-8)) == -1
    assert lib.sl3_step_guarded(None, None, ctypes.c_double(1.0), 0, None, None, None, None,
                                ctypes.c_double(0.1), None) == -1
    count = ctypes.c_int64()
    assert lib.sl3_spectral_rows(None, 4, ctypes.byref(count)) == -1


def test_tuning_override_rejects_unknown_tile_shapes_without_a_device():
    """sl3_tuning_set checks its shapes before anything else: status -1 for
    a shape the kernels are not built for; zeros (tune) are accepted."""
    import sl3dpp_amd

    lib = sl3dpp_amd.lib
    assert lib.sl3_tuning_set(48, 0, 0) == -1
    assert lib.sl3_tuning_set(0, 100, 0) == -1
    assert lib.sl3_tuning_set(16, 33, 0) == -1
    assert lib.sl3_tuning_set(0, 0, 12) == -1
    assert lib.sl3_tuning_set(0, 0, 0) == 0
