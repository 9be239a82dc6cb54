"""Host-side check of the device-memory arena's bookkeeping (best-fit
allocation, splitting, coalescing, fit prediction): tests/native/
arena_check.cpp is compiled with g++ against the library's own
dn_arena_pool.cpp and run; no device is involved."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "denise-black-edition_amd", "csrc")


def test_arena_bookkeeping(tmp_path):
    exe = tmp_path / "arena_check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", CSRC, os.path.join(HERE, "native", "arena_check.cpp"),
                    os.path.join(CSRC, "dn_arena_pool.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr + out.stdout
    assert "arena bookkeeping ok" in out.stdout
