"""CPU oracle for the N-Queens counting path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only as the checker. It wraps liboracle.so (plain C, built
from nq_oracle.c by the Makefile beside it), a restatement of the reference's
counting loop (GPU_NQueens.cu:338-452), pool generator (n_queens.cu:27-219)
and pool file layout (nq_utils.cu:127-259). See nq_oracle.h for what pins
each function.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

U64P = ctypes.POINTER(ctypes.c_uint64)
U8P = ctypes.POINTER(ctypes.c_uint8)


def build():
    """Compile liboracle.so with gcc (the oracle's own Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_count_prefix.argtypes = [ctypes.c_uint, ctypes.c_uint, U8P, U64P, U64P]
        L.orc_count_prefix.restype = ctypes.c_int
        L.orc_count_pool.argtypes = [ctypes.c_uint, ctypes.c_uint, U8P, ctypes.c_uint64,
                                     ctypes.c_int, ctypes.c_uint, U64P, U64P, U64P, U64P]
        L.orc_count_pool.restype = ctypes.c_int
        L.orc_pool_count.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_int, U64P]
        L.orc_pool_count.restype = ctypes.c_uint64
        L.orc_pool_generate.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_int, U8P,
                                        ctypes.c_uint64]
        L.orc_pool_generate.restype = ctypes.c_uint64
        L.orc_lock_depth.argtypes = [ctypes.c_uint, ctypes.c_uint64, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_uint), U64P]
        L.orc_lock_depth.restype = ctypes.c_int
        L.orc_level_counts.argtypes = [ctypes.c_uint, ctypes.c_uint, U64P]
        L.orc_level_counts.restype = None
        L.orc_nqsf_write.argtypes = [ctypes.c_uint, ctypes.c_uint, U8P, ctypes.c_uint64,
                                     ctypes.c_uint64, U8P, ctypes.c_uint64]
        L.orc_nqsf_write.restype = ctypes.c_uint64
        L.orc_sym_count.argtypes = [ctypes.c_uint, U64P, U64P, U64P]
        L.orc_sym_count.restype = ctypes.c_int
        L.orc_sym_brute.argtypes = [ctypes.c_uint, U64P, U64P]
        L.orc_sym_brute.restype = ctypes.c_int
        L.orc_fullsweep_count.argtypes = [ctypes.c_uint, ctypes.c_uint, U64P, U64P]
        L.orc_fullsweep_count.restype = ctypes.c_int
        L.orc_fullsweep_subtree.argtypes = [ctypes.c_uint, ctypes.c_uint32, ctypes.c_uint32,
                                            ctypes.c_uint64, ctypes.c_uint64, U64P, U64P]
        L.orc_fullsweep_subtree.restype = ctypes.c_int
        L.orc_fullsweep_children.argtypes = [ctypes.c_uint, ctypes.c_uint32, ctypes.c_uint32,
                                             ctypes.c_uint64, ctypes.c_uint64, U64P, ctypes.c_uint,
                                             U64P]
        L.orc_fullsweep_children.restype = ctypes.c_int
        _lib = L
    return _lib


def _u8(a):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    return a, a.ctypes.data_as(U8P)


def count_prefix(n, cols):
    """(solutions, nodes below the prefix) for one prefix; ValueError if invalid."""
    arr, p = _u8(list(cols) if len(cols) else [0])
    s, k = ctypes.c_uint64(), ctypes.c_uint64()
    rc = lib().orc_count_prefix(n, len(cols), p, ctypes.byref(s), ctypes.byref(k))
    if rc != 0:
        raise ValueError("invalid prefix")
    return s.value, k.value


def count_pool(n, depth, pool, mirror=False, threads=None, per_prefix=False):
    """Total (mirror-weighted) solutions and nodes below a pool of prefixes.

    pool: uint8 array of shape (count, depth). With per_prefix=True also
    returns unweighted per-prefix solution and node arrays.
    """
    pool = np.ascontiguousarray(pool, dtype=np.uint8).reshape(-1, depth) if depth else \
        np.zeros((len(pool) if hasattr(pool, "__len__") else 0, 0), np.uint8)
    count = pool.shape[0]
    threads = threads or os.cpu_count() or 1
    tot, nd = ctypes.c_uint64(), ctypes.c_uint64()
    ps = pn = None
    ps_p = pn_p = None
    if per_prefix:
        ps = np.zeros(count, np.uint64)
        pn = np.zeros(count, np.uint64)
        ps_p, pn_p = ps.ctypes.data_as(U64P), pn.ctypes.data_as(U64P)
    buf = pool if pool.size else np.zeros(1, np.uint8)
    rc = lib().orc_count_pool(n, depth, buf.ctypes.data_as(U8P), count, int(bool(mirror)),
                              threads, ctypes.byref(tot), ctypes.byref(nd), ps_p, pn_p)
    if rc != 0:
        raise ValueError("invalid prefix in pool")
    if per_prefix:
        return tot.value, nd.value, ps, pn
    return tot.value, nd.value


def pool_count(n, lock, mirror=False, per_col=False):
    pc = np.zeros(max(n, 1), np.uint64)
    tot = lib().orc_pool_count(n, lock, int(bool(mirror)), pc.ctypes.data_as(U64P))
    return (tot, pc) if per_col else tot


def pool_generate(n, lock, mirror=False):
    cnt = pool_count(n, lock, mirror)
    out = np.zeros((max(cnt, 1), lock), np.uint8)
    got = lib().orc_pool_generate(n, lock, int(bool(mirror)), out.ctypes.data_as(U8P), cnt)
    assert got == cnt
    return out[:cnt]


def lock_depth(n, k, mirror=True):
    lk, cnt = ctypes.c_uint(), ctypes.c_uint64()
    rc = lib().orc_lock_depth(n, k, int(bool(mirror)), ctypes.byref(lk), ctypes.byref(cnt))
    if rc != 0:
        raise ValueError("lock-depth search fails for this request")
    return lk.value, cnt.value


def level_counts(n, max_k=None):
    """Valid partial placements per level k = 0..max_k (default n)."""
    max_k = n if max_k is None else min(max_k, n)
    lv = np.zeros(n + 1, np.uint64)
    lib().orc_level_counts(n, max_k, lv.ctypes.data_as(U64P))
    return [int(x) for x in lv[:max_k + 1]]


def nqsf_bytes(n, lock, pool, flags=3):
    pool = np.ascontiguousarray(pool, dtype=np.uint8).reshape(-1, lock)
    cap = 22 + 9 + pool.size
    buf = np.zeros(cap, np.uint8)
    src = pool if pool.size else np.zeros(1, np.uint8)
    w = lib().orc_nqsf_write(n, lock, src.ctypes.data_as(U8P), pool.shape[0], flags,
                             buf.ctypes.data_as(U8P), cap)
    assert w > 0
    return bytes(buf[:w])


def sym_count(n):
    """Symmetry-reduced count (nq_symmetry.c): (unique classes, total
    solutions, placements of the reduced search)."""
    u, t, nd = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    rc = lib().orc_sym_count(n, ctypes.byref(u), ctypes.byref(t), ctypes.byref(nd))
    if rc != 0:
        raise ValueError(f"orc_sym_count({n}) = {rc}")
    return u.value, t.value, nd.value


def fullsweep_count(n, threads=1):
    """Full DoubleSweep count (nq_fullsweep.c): (solutions, placements) of the
    mirror-halved search."""
    s, d = ctypes.c_uint64(), ctypes.c_uint64()
    rc = lib().orc_fullsweep_count(n, threads, ctypes.byref(s), ctypes.byref(d))
    if rc != 0:
        raise ValueError(f"orc_fullsweep_count({n}) = {rc}")
    return s.value, d.value


def fullsweep_subtree(n, state):
    """(solutions, placements) below one state (P, C, D1, D2), weight 1."""
    s, d = ctypes.c_uint64(), ctypes.c_uint64()
    rc = lib().orc_fullsweep_subtree(n, *state, ctypes.byref(s), ctypes.byref(d))
    if rc != 0:
        raise ValueError(f"orc_fullsweep_subtree({n}) = {rc}")
    return s.value, d.value


def fullsweep_children(n, state, cap=32):
    """(children as (P, C, D1, D2) tuples, propagated placements) of one state."""
    out = np.zeros(4 * cap, dtype=np.uint64)
    d = ctypes.c_uint64()
    k = lib().orc_fullsweep_children(n, *state, out.ctypes.data_as(U64P), cap, ctypes.byref(d))
    if k < 0:
        raise ValueError(f"orc_fullsweep_children({n}) = {k}")
    rows = out[: 4 * min(k, cap)].reshape(-1, 4)
    return [tuple(int(v) for v in r) for r in rows], d.value


def sym_brute(n):
    """Brute force over every solution's 8 images: (unique, total)."""
    u, t = ctypes.c_uint64(), ctypes.c_uint64()
    rc = lib().orc_sym_brute(n, ctypes.byref(u), ctypes.byref(t))
    if rc != 0:
        raise ValueError(f"orc_sym_brute({n}) = {rc}")
    return u.value, t.value
