"""The C++ interface (include/sl3dpp_amd.hpp, the reference's
Viscoelastic3D names and exception types over the C ABI) from a C++ caller:
tests/cpp/step_api.cpp, built by tests/cpp/Makefile. On the CPU: an Earth
from a layered spec, and a bad spec raising std::invalid_argument."""

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "step_api")


def test_cpp_interface_on_the_host():
    # the test program is host code only (g++), built like the oracle
    res = subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "cpp")],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    run = subprocess.run([BIN, "cpu"], capture_output=True, text=True, timeout=120)
    print(run.stdout)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "[cpp] ok" in run.stdout
