"""The C++ API (include/tamm/tamm.hpp) used the way TAMM's own tests use TAMM: the program
tests/cpp/test_api.cpp restates Test_Ops / Test_Mult_Ops checks and a spin-blocked
ladder + ring against a dense loop, all through Scheduler(...).execute(ExecutionHW::GPU)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_api")


def test_cpp_api_program_is_built_and_linked():
    if not os.path.exists(BIN):
        from tamm_amd import build
        build.build_cpp_tests()
    out = subprocess.run(["ldd", BIN], capture_output=True, text=True).stdout
    assert "libtamm_mi355x.so" in out and "not found" not in out.split("libtamm_mi355x.so")[1].split("\n")[0]


@pytest.mark.gpu
def test_cpp_api_program_passes():
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all passed" in r.stdout
    assert "norm 262500000" in r.stdout
