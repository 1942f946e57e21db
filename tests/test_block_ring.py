"""CPU test of the pinned block ring (spatial/block_ring.hpp): compiled with
g++ and run with a fake pipeline, no GPU."""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)


def test_block_ring_order_reuse_and_failure():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "test_block_ring")
        subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-pthread",
                        "-I", os.path.join(REPO, "cuda-spatial-filtering_amd", "include"),
                        "-I", os.path.join(REPO, "include"),
                        os.path.join(HERE, "cpp", "test_block_ring.cpp"), "-o", exe], check=True)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "0 failures" in r.stdout
