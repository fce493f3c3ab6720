"""Host check of the scratch bookkeeping behind the plain entry points
(pymdp_amd/csrc/scratch_book.hpp): buffers handed to a call are retired, not
freed, when a larger call grows them, so a graph captured earlier keeps valid
addresses; different streams never share a buffer; nothing is allocated
under stream capture.  Built with g++ and run on the CPU."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_scratch_book_never_frees_live_buffers(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "scratch_book_check"
    subprocess.run([gxx, "-std=c++17", "-O1", "-Wall", "-I", str(ROOT / "pymdp_amd" / "csrc"),
                    str(ROOT / "tests" / "native" / "scratch_book_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
