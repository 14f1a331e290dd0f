"""The C++ SMAX::Interface (HIP platform) through its own test runner
(tests/cpp/run_tests.cpp): host-side API behaviour on the CPU, the
reference's known answers on the GPU."""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

RUNNER = Path(__file__).resolve().parent / "cpp" / "_build" / "run_tests"


def run(group: str) -> str:
    if not RUNNER.exists():
        pytest.fail(f"{RUNNER} not built (run __graft_entry__.build())")
    p = subprocess.run([str(RUNNER), group], capture_output=True, text=True,
                       timeout=300, cwd=RUNNER.parent)
    assert p.returncode == 0, p.stdout + p.stderr
    return p.stdout


def test_interface_host_behaviour():
    out = run("host")
    assert "0 failed" in out


@pytest.mark.gpu
def test_interface_hip_platform_kats():
    out = run("gpu")
    assert "0 failed" in out and "FAILED" not in out
