"""The SBA service with the MI355X bundler on the GPU: tests/cpp/service_gpu.cpp (a plain g++ program of the
reference's API) builds a map of the reference's production window (7 key frames of which 3 fixed, 7 LM
iterations), runs run_sba / SbaService (libs/pipelines/service_sba.h:72-180, 301-348) with
SchurComplementBundlerGpu, and checks the solve, the write-back of the solved state into the map, the
bit-for-bit agreement with a direct solve of the assembled problem, the asynchronous worker and the winsorizer's
second solve. The direct solve is the one the drop-in and parity tests compare with the FP64 oracle."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "service_gpu")


def test_service_solves_and_writes_back_the_map():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (__graft_entry__.build() builds it)")
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().splitlines()[-1] == "OK (0 failures)"
