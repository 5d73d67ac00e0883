"""The C-ABI library loads and exports every entry point include/nqm.h
declares. No compute calls: this file runs without a GPU."""
import ctypes
import os
import re

import pytest

import nqm
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "nqm.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nqm_\w+)\s*\(", text)))


def test_library_is_built():
    assert os.path.exists(nqm.LIB_PATH), "run gpu-n-queens_amd/build.py"


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(nqm.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    # every header entry point has a ctypes signature in the Python binding
    assert set(declared_functions()) <= set(nqm.exported_symbols())


def test_strerror_and_state_size_need_no_device():
    assert nqm.strerror(nqm.OK) == "ok"
    assert nqm.strerror(nqm.EPREFIX) == "invalid prefix"
    # sizeof(nq_state_t), n_queens.cuh:100-113: 40 bytes for 12 <= n <= 19
    for n in range(12, 20):
        assert nqm.state_size(n) == 40
    assert nqm.state_size(20) == 48
    assert nqm.state_size(11) == 32
    assert nqm.state_size(0) == 0 and nqm.state_size(32) == 0
