"""CPU tests of the header-only SBA service (include/sba/service_sba.h): the C++ program tests/cpp/service_test.cpp
with a recording bundler (problem assembly from a submap, concurrent triangulation, fixed key frames, write-back,
the worker thread, the winsorizer between two solves), and the same program built with the thread and
address/undefined-behaviour sanitizers."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")


def build(target):
    subprocess.run(["make", "-s", "-C", CPP, target], check=True)
    return os.path.join(CPP, target)


def test_service_cases():
    exe = build("service_test")
    out = subprocess.run([exe, "1000"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "OK (0 failures)" in out.stdout


@pytest.mark.parametrize("target", ["service_test_tsan", "service_test_asan"])
def test_service_cases_under_sanitizers(target):
    exe = build(target)
    out = subprocess.run([exe, "50"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "OK (0 failures)" in out.stdout


@pytest.mark.parametrize("target", ["runtime_host_check", "runtime_host_check_tsan"])
def test_host_runtime_binds_every_stream_inside_its_block(target):
    # the host runtime driven without a GPU (tests/cpp/runtime_host_check.cpp): every stream each kernel launch
    # receives is bound, inside a live allocation, and apart from the others; sba_solve_noalloc allocates nothing;
    # packing on several host threads gives the same streams (under the address and the thread sanitizers)
    exe = build(target)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().splitlines()[-1].startswith("OK"), out.stdout
