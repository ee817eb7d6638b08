"""The C++ wrapper (include/hdcx_path_tracer.hpp) compiled with g++ and run
like the reference's integrator test executable."""

import subprocess
from pathlib import Path

import pytest

import hdcodex_amd as hx

ROOT = Path(__file__).resolve().parent.parent


def _build(tmp_path: Path) -> Path:
    exe = tmp_path / "wrapper_test"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", str(ROOT / "include"), str(ROOT / "tests" / "cpp" / "wrapper_test.cpp"),
                    "-L", str(hx.LIB_PATH.parent), "-lhdcx", f"-Wl,-rpath,{hx.LIB_PATH.parent}", "-o", str(exe)],
                   check=True, capture_output=True)
    return exe


def test_wrapper_compiles_and_reports_missing_device(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    # 3 = constructor threw std::runtime_error (no device), 0 = full pass.
    assert r.returncode in (0, 3), r.stdout + r.stderr


@pytest.mark.gpu
def test_wrapper_renders_on_the_gpu(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "wrapper ok" in r.stdout
