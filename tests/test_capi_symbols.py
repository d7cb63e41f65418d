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
