"""include/gpe_queue.h between two threads, in the reference driver's pattern
(test_vortex/rte_vortex.c:64-89, 310-370; dev/gpe_queue.h:67-139): every frame
reaches the writer once, in order and unmixed, for ring sizes 1 to 5."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def binary(tmp_path_factory):
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("no C++ compiler")
    out = str(tmp_path_factory.mktemp("queue") / "queue_check")
    subprocess.check_call([cxx, "-std=c++17", "-O1", "-pthread", "-I" + os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "drivers", "queue_check.cpp"), "-o", out])
    return out


@pytest.mark.parametrize("qsize", [1, 2, 5])
def test_frames_arrive_in_order_and_unmixed(binary, qsize):
    r = subprocess.run([binary, "150", str(qsize)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout
    assert "frames 150, corrupted 0" in r.stdout
