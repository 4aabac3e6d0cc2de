"""The device heap emulation compiled for the host and checked against
libstdc++'s std::priority_queue with the reference comparator
(reorder/swappair.h:112-120): pop orders of 300,000 random heaps with ties,
under AddressSanitizer so that any read or write outside a heap's 33 words
fails.  Needs hipcc (the HIP headers), not a GPU."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "vnm-graph-reordering_amd", "csrc")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc is not installed")
def test_heap_emulation_matches_libstdcxx_priority_queue(tmp_path):
    exe = str(tmp_path / "heap_check")
    subprocess.run([HIPCC, "-O1", "-g", "-std=c++17", "-Xarch_host", "-fsanitize=address",
                    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
                    os.path.join(HERE, "heap_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    out = subprocess.run([exe, "300000"], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip() == "cases 300000 bad 0", out.stdout
