"""The header-only C++ drop-in classes (include/rdu_rendu.hpp) compile against Rendu-shaped types,
link against librdu.so and answer a query (on a GPU) or log the library's error (without one)."""
import os
import subprocess

import pytest

from conftest import GPU, ROOT


def _build(tmp_path):
    exe = os.path.join(str(tmp_path), "dropin")
    subprocess.run(["g++", "-std=c++14", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "dropin_example.cpp"), "-L", os.path.join(ROOT, "rendu_amd"),
                    "-lrdu", f"-Wl,-rpath,{os.path.join(ROOT, 'rendu_amd')}", "-o", exe], check=True)
    return exe


@pytest.mark.skipif(GPU, reason="checks the no-device path")
def test_dropin_compiles_and_logs_without_gpu(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 3 and "no HIP device" in r.stdout


@pytest.mark.gpu
def test_dropin_queries_on_gpu(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    # Quad z = 0, ray from (0.5, -0.25, 1) straight down: t = 1, first triangle (local id 0).
    assert "hit=1 dist=1.000000" in r.stdout and "local=0" in r.stdout and "any=1" in r.stdout
