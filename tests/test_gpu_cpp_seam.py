"""Runs the C++ seam tests (tests/cpp/test_pipelines.cpp) of the
GPUPipeline-compatible classes on the GPU: the reference's corr-beam and
adaptive golden tests through the C++ API (built by __graft_entry__.build())."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "_build", "test_pipelines")


def test_cpp_pipeline_seam():
    assert os.path.exists(BIN), "C++ seam tests are not built (run __graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=110)
    print(r.stdout[-4000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout[-4000:]
    assert "0 failures" in r.stdout
