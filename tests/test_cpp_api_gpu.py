"""The C++ interface (include/sl3dpp_amd.hpp) stepping on the device, from
the C++ test program tests/cpp/step_api.cpp (built with the library): Step
returns a new state and leaves its argument unchanged, two steps by Step
equal two by StepInPlace bit for bit, the solutions hold every degree, a
negative dt and a degree-1 tide raise std::invalid_argument (the state
unchanged), and a load channel's own exception comes out of Step."""

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "step_api")


def test_cpp_interface_steps():
    assert os.path.exists(BIN), "tests/cpp/bin/step_api is not built (run __graft_entry__.build())"
    print("[cpp] step_api gpu", flush=True)
    run = subprocess.run([BIN, "gpu"], capture_output=True, text=True, timeout=300)
    print(run.stdout)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "[cpp] ok" in run.stdout
