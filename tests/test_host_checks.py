"""Host runs of device search code (CPU only): TieSearch, the bottom-up search
behind the symmetric drain's tie records (gpu-n-queens_amd/csrc/k4_ties.h),
against the tie boards themselves on N=10..13 (tests/host_checks/)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "host_checks")

pytestmark = pytest.mark.skipif(not os.path.exists("/opt/rocm/bin/hipcc") or shutil.which("make") is None,
                                reason="needs hipcc and make")


def test_tie_search_finds_every_recorded_board():
    subprocess.run(["make", "-s", "-C", HERE], check=True, timeout=300)
    r = subprocess.run([os.path.join(HERE, "_build", "ties_check"), "10", "13"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("ok")
