import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("denise-black-edition_amd", "oracle", "tests"):
    path = os.path.join(ROOT, sub)
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD MI355X GPU (runs the HIP library)")


def pytest_sessionstart(session):
    """Build the CPU oracle and the libraries when a fresh checkout lacks them
    (they are build products, not tracked files)."""
    jobs = str(min(8, os.cpu_count() or 4))
    need = [("oracle", "liboracle_f32.so"), ("denise-black-edition_amd", "libdnhip.so")]
    for sub, lib in need:
        d = os.path.join(ROOT, sub)
        if not os.path.exists(os.path.join(d, lib)):
            subprocess.run(["make", "-C", d, "-j", jobs], check=True)
