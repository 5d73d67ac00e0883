"""Host sanitizer runs (CPU only): the .nqsf parser and state converters
(gpu-n-queens_amd/csrc/nqm_files.hip) on seeded damaged files, and the CPU
oracle on small boards against known answers, both built with
AddressSanitizer and UndefinedBehaviorSanitizer by tests/sanitize/Makefile."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None,
                                reason="needs g++ and make")


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-s", "-C", HERE], check=True, timeout=300)
    return os.path.join(HERE, "_build")


def test_nqsf_parser_under_asan_ubsan(built, tmp_path):
    r = subprocess.run([os.path.join(built, "nqsf_fuzz"), str(tmp_path), "4000"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "rejected" in r.stdout


def test_oracle_under_asan_ubsan(built):
    r = subprocess.run([os.path.join(built, "oracle_check")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
