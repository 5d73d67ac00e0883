"""Python binding of libnqm.so (the C ABI declared in include/nqm.h).

A thin ctypes layer for tests, the bench and scripting: every call goes
through the C entry points, which run the HIP kernels. There is no CPU
fallback. If libnqm.so is missing or no wave64 GPU is present, the calls
raise.

Usage:
    import sys; sys.path.insert(0, "<repo>/gpu-n-queens_amd"); import nqm
    with nqm.Context(0) as ctx:
        total = ctx.count_pool(16, 5, pool)          # pool: uint8 (count, 5)
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# NQM_LIB names another build of the same library (for comparing builds).
LIB_PATH = os.environ.get("NQM_LIB") or os.path.join(_HERE, "libnqm.so")

OK = 0
EINVAL = 1
EPREFIX = 2
ENOMEM = 3
EDEVICE = 4
ERANGE = 5
EIO = 6
EFORMAT = 7
EFLAGS = 8
EUNSUPPORTED = 9

MIRROR_WEIGHT = 0x1

U8P = ctypes.POINTER(ctypes.c_uint8)
U64P = ctypes.POINTER(ctypes.c_uint64)
VP = ctypes.c_void_p


class NqmError(RuntimeError):
    def __init__(self, status, detail=""):
        self.status = status
        msg = f"nqm status {status}"
        if _lib is not None:
            msg += f" ({_lib.nqm_strerror(status).decode()})"
        if detail:
            msg += f": {detail}"
        super().__init__(msg)


class Stats(ctypes.Structure):
    _fields_ = [("total", ctypes.c_uint64), ("nodes", ctypes.c_uint64),
                ("kernel_ms", ctypes.c_double), ("wall_ms", ctypes.c_double),
                ("launches", ctypes.c_uint32), ("blocks", ctypes.c_uint32),
                ("search_ms", ctypes.c_double), ("rounds", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("lanes", ctypes.c_uint64), ("parked_max", ctypes.c_uint64),
                ("handed_back_max", ctypes.c_uint64), ("hand_back_room", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Solve(ctypes.Structure):
    """nqm_solve: board size, pre-placed queens and their columns."""
    _fields_ = [("n", ctypes.c_uint32), ("k", ctypes.c_uint32), ("prefix", ctypes.c_uint8 * 32)]


class Tune(ctypes.Structure):
    """nqm_tune (zero = automatic for every field)."""
    _fields_ = [("depth", ctypes.c_uint32), ("waves_per_cu", ctypes.c_uint32),
                ("threads", ctypes.c_uint32), ("refill_idle", ctypes.c_uint32),
                ("budget", ctypes.c_uint64), ("flags", ctypes.c_uint32),
                ("stack_rows", ctypes.c_uint32), ("model", ctypes.c_void_p),
                ("split_log2", ctypes.c_float), ("split_rounds", ctypes.c_uint32)]


TUNE_NO_MIRROR = 0x1
TUNE_NO_DONATE = 0x2
TUNE_GRAPH = 0x4
TUNE_NO_ORDER = 0x8


def make_solves(solves):
    arr = (Solve * len(solves))()
    for i, (n, prefix) in enumerate(solves):
        arr[i].n = n
        arr[i].k = len(prefix)
        for r, c in enumerate(prefix):
            arr[i].prefix[r] = c
    return arr


def make_tune(**kw):
    t = Tune()
    for k, v in kw.items():
        if k == "model" and v is not None and not isinstance(v, int):
            v = v.handle
        setattr(t, k, v)
    return t


class TuneSpace(ctypes.Structure):
    """nqm_tune_space: candidate lists (zero-terminated, up to 8 each)."""
    _fields_ = [("depths", ctypes.c_uint32 * 8), ("waves_per_cu", ctypes.c_uint32 * 8),
                ("stack_rows", ctypes.c_uint32 * 8), ("repeats", ctypes.c_uint32),
                ("base_flags", ctypes.c_uint32)]


class TuneResult(ctypes.Structure):
    _fields_ = [("depth", ctypes.c_uint32), ("waves_per_cu", ctypes.c_uint32),
                ("stack_rows", ctypes.c_uint32), ("pad", ctypes.c_uint32), ("ms", ctypes.c_double)]


class ServerCfg(ctypes.Structure):
    """nqm_server_cfg (zero = default for every field)."""
    _fields_ = [("max_active", ctypes.c_uint32), ("depth", ctypes.c_uint32),
                ("budget", ctypes.c_uint64), ("queue_bytes", ctypes.c_uint64),
                ("flags", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class Result(ctypes.Structure):
    """nqm_result: one finished request."""
    _fields_ = [("tag", ctypes.c_uint64), ("count", ctypes.c_uint64), ("nodes", ctypes.c_uint64),
                ("rounds", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("admit_ms", ctypes.c_double), ("done_ms", ctypes.c_double)]


class GpuConfig(ctypes.Structure):
    """gpu_config_t of the reference (GPU_NQueens.cuh:8-11)."""
    _fields_ = [("device_id", ctypes.c_uint), ("async_", ctypes.c_int)]


_lib = None

# Exported symbols and their (argtypes, restype); the not-gpu tests check that
# the library exports each of them.
_SIGNATURES = {
    "nqm_create": ([ctypes.c_int, ctypes.POINTER(VP)], ctypes.c_int),
    "nqm_destroy": ([VP], None),
    "nqm_strerror": ([ctypes.c_int], ctypes.c_char_p),
    "nqm_last_error": ([VP], ctypes.c_char_p),
    "nqm_device_info": ([VP] + [ctypes.POINTER(ctypes.c_uint32)] * 5, ctypes.c_int),
    "nqm_count_pool": ([VP, ctypes.c_uint32, ctypes.c_uint32, U8P, ctypes.c_uint64,
                        ctypes.c_uint32, U64P], ctypes.c_int),
    "nqm_count_pool_device": ([VP, ctypes.c_uint32, ctypes.c_uint32, VP, ctypes.c_uint64,
                               ctypes.c_uint32, VP, VP, ctypes.POINTER(Stats)], ctypes.c_int),
    "nqm_count_pool_device_static": ([VP, ctypes.c_uint32, ctypes.c_uint32, VP, ctypes.c_uint64,
                                      ctypes.c_uint32, VP], ctypes.c_int),
    "nqm_solver_driver": ([ctypes.c_uint32, VP, ctypes.c_uint32, ctypes.c_uint,
                           ctypes.POINTER(GpuConfig), ctypes.c_uint, U64P], ctypes.c_int),
    "nqm_state_size": ([ctypes.c_uint32], ctypes.c_size_t),
    "nqm_solve_batch": ([VP, ctypes.POINTER(Solve), ctypes.c_uint32, ctypes.POINTER(Tune), U64P,
                         U64P, ctypes.POINTER(Stats)], ctypes.c_int),
    "nqm_count_symmetric": ([VP, ctypes.c_uint32, U64P, U64P, ctypes.POINTER(Stats)], ctypes.c_int),
    "nqm_count_symmetric_path": ([VP, ctypes.c_uint32, ctypes.c_uint32, U64P, U64P, ctypes.POINTER(Stats)],
                                 ctypes.c_int),
    "nqm_count_fullsweep": ([VP, ctypes.c_uint32, U64P, ctypes.POINTER(Stats)], ctypes.c_int),
    "nqm_count_fullsweep_ex": ([VP, ctypes.c_uint32, ctypes.c_uint32, U64P, ctypes.POINTER(Stats)],
                               ctypes.c_int),
    "nqm_pool_count": ([VP, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, U64P, U64P],
                       ctypes.c_int),
    "nqm_pool_generate": ([VP, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, U8P,
                           ctypes.c_uint64, U64P], ctypes.c_int),
    "nqm_pool_generate_device": ([VP, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                  ctypes.POINTER(VP), U64P], ctypes.c_int),
    "nqm_choose_lock": ([VP, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32,
                         ctypes.POINTER(ctypes.c_uint32), U64P], ctypes.c_int),
    "nqm_nqsf_write": ([ctypes.c_char_p, ctypes.c_uint32, ctypes.c_uint32, U8P, ctypes.c_uint64,
                        ctypes.c_uint64], ctypes.c_int),
    "nqm_nqsf_read": ([ctypes.c_char_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(U8P), U64P,
                       ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32), U64P],
                      ctypes.c_int),
    "nqm_free": ([VP], None),
    "nqm_states_to_pool": ([ctypes.c_uint32, VP, ctypes.c_uint64, ctypes.c_uint32, U8P], ctypes.c_int),
    "nqm_pool_to_states": ([ctypes.c_uint32, ctypes.c_uint32, U8P, ctypes.c_uint64, VP], ctypes.c_int),
    "nqm_mlp_param_count": ([], ctypes.c_uint32),
    "nqm_mlp_feature_width": ([], ctypes.c_uint32),
    "nqm_mlp_create": ([VP, ctypes.c_uint64, ctypes.POINTER(VP)], ctypes.c_int),
    "nqm_mlp_destroy": ([VP], None),
    "nqm_mlp_set_params": ([VP, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "nqm_mlp_get_params": ([VP, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "nqm_mlp_features_pool": ([VP, ctypes.c_uint32, ctypes.c_uint32, VP, ctypes.c_uint64, VP],
                              ctypes.c_int),
    "nqm_mlp_predict": ([VP, VP, ctypes.c_uint32, VP], ctypes.c_int),
    "nqm_mlp_gradient": ([VP, VP, VP, ctypes.c_uint32, ctypes.POINTER(ctypes.c_float),
                          ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "nqm_mlp_train": ([VP, VP, VP, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                       ctypes.c_uint64, ctypes.c_float, ctypes.POINTER(ctypes.c_float),
                       ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "nqm_autotune": ([VP, ctypes.POINTER(Solve), ctypes.c_uint32, ctypes.POINTER(TuneSpace),
                      ctypes.POINTER(Tune), ctypes.POINTER(TuneResult), ctypes.c_uint32,
                      ctypes.POINTER(ctypes.c_uint32)], ctypes.c_int),
    "nqm_server_create": ([VP, ctypes.POINTER(ServerCfg), ctypes.POINTER(VP)], ctypes.c_int),
    "nqm_server_destroy": ([VP], None),
    "nqm_server_submit": ([VP, ctypes.POINTER(Solve), U64P, ctypes.c_uint32,
                           ctypes.POINTER(ctypes.c_uint32)], ctypes.c_int),
    "nqm_server_step": ([VP, ctypes.POINTER(ctypes.c_uint32)], ctypes.c_int),
    "nqm_server_poll": ([VP, ctypes.POINTER(Result), ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)],
                        ctypes.c_int),
    "nqm_server_info": ([VP, ctypes.POINTER(ctypes.c_uint32), U64P, U64P], ctypes.c_int),
    "nqm_server_round_stats": ([VP, ctypes.POINTER(Stats)], ctypes.c_int),
    "nqm_server_clock_ms": ([VP], ctypes.c_double),
    "nqm_dev_alloc": ([VP, ctypes.c_size_t, ctypes.POINTER(VP)], ctypes.c_int),
    "nqm_dev_free": ([VP, VP], ctypes.c_int),
    "nqm_memcpy_h2d": ([VP, VP, VP, ctypes.c_size_t], ctypes.c_int),
    "nqm_dev_copy": ([VP, VP, VP, ctypes.c_size_t, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "nqm_memcpy_d2h": ([VP, VP, VP, ctypes.c_size_t], ctypes.c_int),
}


def load():
    """Load libnqm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run gpu-n-queens_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGNATURES)


def strerror(status):
    return load().nqm_strerror(status).decode()


def state_size(n):
    return int(load().nqm_state_size(n))


def _pool_array(pool, depth):
    a = np.ascontiguousarray(pool, dtype=np.uint8)
    if depth:
        a = a.reshape(-1, depth)
    return a


class DeviceBuffer:
    """Device memory owned through the library (nqm_dev_alloc)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = VP()
        ctx._check(ctx.lib.nqm_dev_alloc(ctx.handle, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, ctx, array):
        a = np.ascontiguousarray(array)
        buf = cls(ctx, max(a.nbytes, 1))
        if a.nbytes:
            ctx._check(ctx.lib.nqm_memcpy_h2d(ctx.handle, buf.ptr, a.ctypes.data, a.nbytes))
        return buf

    def to_host(self, dtype, count):
        out = np.zeros(count, dtype=dtype)
        if out.nbytes:
            self.ctx._check(self.ctx.lib.nqm_memcpy_d2h(self.ctx.handle, out.ctypes.data,
                                                        self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            self.ctx.lib.nqm_dev_free(self.ctx.handle, self.ptr)
            self.ptr = None


class Context:
    """One nqm_ctx (device, stream, events, cached device buffers)."""

    def __init__(self, device=0):
        self.lib = load()
        h = VP()
        rc = self.lib.nqm_create(device, ctypes.byref(h))
        if rc != OK:
            raise NqmError(rc, f"nqm_create(device={device})")
        self.handle = h.value

    def close(self):
        if self.handle:
            self.lib.nqm_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != OK:
            raise NqmError(rc, self.lib.nqm_last_error(self.handle).decode())
        return rc

    def last_error(self):
        return self.lib.nqm_last_error(self.handle).decode()

    def device_info(self):
        vals = [ctypes.c_uint32() for _ in range(5)]
        self._check(self.lib.nqm_device_info(self.handle, *[ctypes.byref(v) for v in vals]))
        keys = ["cus", "wave_size", "clock_khz", "lds_per_block", "lds_per_cu"]
        return {k: v.value for k, v in zip(keys, vals)}

    # -- the counting path -------------------------------------------------
    def count_pool(self, n, depth, pool, flags=0, count=None):
        """Total (weighted) solutions below a host pool (nqm_count_pool)."""
        a = _pool_array(pool, depth)
        cnt = a.shape[0] if count is None else count
        tot = ctypes.c_uint64()
        src = a if a.size else np.zeros(1, np.uint8)
        self._check(self.lib.nqm_count_pool(self.handle, n, depth, src.ctypes.data_as(U8P), cnt,
                                            flags, ctypes.byref(tot)))
        return tot.value

    def count_pool_raw(self, n, depth, pool, flags=0, count=None):
        """Like count_pool but returns the raw status instead of raising."""
        a = _pool_array(pool, depth)
        cnt = a.shape[0] if count is None else count
        tot = ctypes.c_uint64()
        src = a if a.size else np.zeros(1, np.uint8)
        rc = self.lib.nqm_count_pool(self.handle, n, depth, src.ctypes.data_as(U8P), cnt, flags,
                                     ctypes.byref(tot))
        return rc, tot.value

    def count_pool_device(self, n, depth, d_pool, count, flags=0, per_prefix=False):
        """Count a device-resident pool. Returns Stats, plus per-prefix
        (unweighted) solution and node arrays when per_prefix is set."""
        st = Stats()
        ps = pn = None
        if per_prefix:
            ps = DeviceBuffer(self, 8 * count)
            pn = DeviceBuffer(self, 8 * count)
        try:
            self._check(self.lib.nqm_count_pool_device(
                self.handle, n, depth, d_pool.ptr if isinstance(d_pool, DeviceBuffer) else d_pool,
                count, flags, ps.ptr if ps else None, pn.ptr if pn else None, ctypes.byref(st)))
            if per_prefix:
                return st, ps.to_host(np.uint64, count), pn.to_host(np.uint64, count)
            return st
        finally:
            if ps:
                ps.free()
            if pn:
                pn.free()

    def count_pool_device_static(self, n, depth, d_pool, count, d_result, flags=0):
        """The allocation-free count (nqm_count_pool_device_static): enqueues
        the count, whose three u64 results (solutions, nodes, invalid-prefix
        flag) land in the device buffer d_result (a DeviceBuffer of at least
        24 bytes, or a device pointer). Nothing is allocated and the call does
        not wait; read d_result back (DeviceBuffer.to_host waits for the
        context's stream)."""
        self._check(self.lib.nqm_count_pool_device_static(
            self.handle, n, depth, d_pool.ptr if isinstance(d_pool, DeviceBuffer) else d_pool, count, flags,
            d_result.ptr if isinstance(d_result, DeviceBuffer) else d_result))


def _solve_batch(ctx, solves, tune=None, raw=False, **tune_kw):
    arr = make_solves(solves)
    t = tune if tune is not None else (make_tune(**tune_kw) if tune_kw else None)
    counts = np.zeros(len(solves), np.uint64)
    nodes = np.zeros(len(solves), np.uint64)
    st = Stats()
    rc = ctx.lib.nqm_solve_batch(ctx.handle, arr, len(solves), ctypes.byref(t) if t else None,
                                 counts.ctypes.data_as(U64P), nodes.ctypes.data_as(U64P),
                                 ctypes.byref(st))
    if raw:
        return rc, counts, nodes, st
    ctx._check(rc)
    return counts, nodes, st


Context.solve_batch = _solve_batch


SYM_AUTO = 0
SYM_THREAD = 1  # one thread per host item (any n)
SYM_DRAIN = 2   # the persistent drain (10 <= n <= 20)


def _count_symmetric(ctx, n, path=None):
    """Symmetry-reduced count of the empty n x n board: (unique classes,
    total solutions, Stats). path: SYM_THREAD / SYM_DRAIN / SYM_AUTO (None:
    nqm_count_symmetric itself)."""
    u, t, st = ctypes.c_uint64(), ctypes.c_uint64(), Stats()
    if path is None:
        rc = ctx.lib.nqm_count_symmetric(ctx.handle, n, ctypes.byref(u), ctypes.byref(t), ctypes.byref(st))
    else:
        rc = ctx.lib.nqm_count_symmetric_path(ctx.handle, n, path, ctypes.byref(u), ctypes.byref(t),
                                              ctypes.byref(st))
    ctx._check(rc)
    return u.value, t.value, st


Context.count_symmetric = _count_symmetric


def _count_fullsweep(ctx, n, target_states=None):
    """Full DoubleSweep count of the empty n x n board (propagation over every
    open row and column, centre-out branching): (total, Stats). stats.nodes
    are the placements of this search. target_states: where the device
    expansion stops (None: nqm_count_fullsweep's default)."""
    t, st = ctypes.c_uint64(), Stats()
    if target_states is None:
        rc = ctx.lib.nqm_count_fullsweep(ctx.handle, n, ctypes.byref(t), ctypes.byref(st))
    else:
        rc = ctx.lib.nqm_count_fullsweep_ex(ctx.handle, n, target_states, ctypes.byref(t), ctypes.byref(st))
    ctx._check(rc)
    return t.value, st


Context.count_fullsweep = _count_fullsweep


def _pool_count(ctx, n, lock, flags=0, per_col=False):
    cnt = ctypes.c_uint64()
    pc = np.zeros(32, np.uint64)
    ctx._check(ctx.lib.nqm_pool_count(ctx.handle, n, lock, flags, ctypes.byref(cnt),
                                      pc.ctypes.data_as(U64P)))
    return (cnt.value, pc[:n]) if per_col else cnt.value


def _pool_generate(ctx, n, lock, flags=0):
    """Host copy of the device-generated pool, shape (count, lock)."""
    d = VP()
    cnt = ctypes.c_uint64()
    ctx._check(ctx.lib.nqm_pool_generate_device(ctx.handle, n, lock, flags, ctypes.byref(d),
                                                ctypes.byref(cnt)))
    out = np.zeros((cnt.value, lock), np.uint8)
    try:
        if out.size:
            ctx._check(ctx.lib.nqm_memcpy_d2h(ctx.handle, out.ctypes.data, d.value, out.size))
    finally:
        ctx.lib.nqm_dev_free(ctx.handle, d.value)
    return out


def _pool_generate_device(ctx, n, lock, flags=0):
    """Device-resident pool: (DeviceBuffer-like object, count)."""
    d = VP()
    cnt = ctypes.c_uint64()
    ctx._check(ctx.lib.nqm_pool_generate_device(ctx.handle, n, lock, flags, ctypes.byref(d),
                                                ctypes.byref(cnt)))
    buf = DeviceBuffer.__new__(DeviceBuffer)
    buf.ctx, buf.nbytes, buf.ptr = ctx, cnt.value * lock, d.value
    return buf, cnt.value


def _choose_lock(ctx, n, how_many, flags=MIRROR_WEIGHT):
    lk, cnt = ctypes.c_uint32(), ctypes.c_uint64()
    ctx._check(ctx.lib.nqm_choose_lock(ctx.handle, n, how_many, flags, ctypes.byref(lk),
                                       ctypes.byref(cnt)))
    return lk.value, cnt.value


Context.pool_count = _pool_count
Context.pool_generate = _pool_generate
Context.pool_generate_device = _pool_generate_device
Context.choose_lock = _choose_lock

NQSF_FLAG_EXPERIMENTAL = 0x1
NQSF_FLAG_HALF = 0x2
NQSF_REFERENCE_FLAGS = NQSF_FLAG_EXPERIMENTAL | NQSF_FLAG_HALF


def nqsf_write(path, n, lock, pool, flags=NQSF_REFERENCE_FLAGS):
    a = _pool_array(pool, lock)
    src = a if a.size else np.zeros(1, np.uint8)
    rc = load().nqm_nqsf_write(os.fsencode(path), n, lock, src.ctypes.data_as(U8P), a.shape[0], flags)
    if rc != OK:
        raise NqmError(rc, f"nqm_nqsf_write({path})")


def nqsf_read(path, expect_n=0, expect_flags=(1 << 64) - 1, raw=False):
    """(n, lock, flags, pool) of a .nqsf file."""
    L = load()
    p = U8P()
    cnt = ctypes.c_uint64()
    n, lk = ctypes.c_uint32(), ctypes.c_uint32()
    fl = ctypes.c_uint64()
    rc = L.nqm_nqsf_read(os.fsencode(path), expect_n, expect_flags, ctypes.byref(p), ctypes.byref(cnt),
                         ctypes.byref(n), ctypes.byref(lk), ctypes.byref(fl))
    if rc != OK:
        if raw:
            return rc, None
        raise NqmError(rc, f"nqm_nqsf_read({path})")
    try:
        pool = np.ctypeslib.as_array(p, shape=(cnt.value * lk.value,)).copy() if cnt.value * lk.value \
            else np.zeros(0, np.uint8)
    finally:
        L.nqm_free(ctypes.cast(p, VP))
    out = (n.value, lk.value, fl.value, pool.reshape(cnt.value, lk.value) if lk.value else pool)
    return (OK, out) if raw else out


def states_to_pool(n, states, lock):
    L = load()
    states = np.ascontiguousarray(states, dtype=np.uint8)
    cnt = states.shape[0]
    out = np.zeros((cnt, lock), np.uint8)
    rc = L.nqm_states_to_pool(n, states.ctypes.data, cnt, lock,
                              (out if out.size else np.zeros(1, np.uint8)).ctypes.data_as(U8P))
    if rc != OK:
        raise NqmError(rc, "nqm_states_to_pool")
    return out


def pool_to_states(n, lock, pool):
    L = load()
    a = _pool_array(pool, lock)
    out = np.zeros((a.shape[0], state_size(n)), np.uint8)
    src = a if a.size else np.zeros(1, np.uint8)
    rc = L.nqm_pool_to_states(n, lock, src.ctypes.data_as(U8P), a.shape[0], out.ctypes.data)
    if rc != OK:
        raise NqmError(rc, "nqm_pool_to_states")
    return out


def states_from_pool(n, lock, pool):
    """Build an array in the reference's nq_state_t layout (n_queens.cuh:100-113)
    from a compact pool, as the reference's loader does (nq_utils.cu:244-251)."""
    pool = _pool_array(pool, lock)
    sz = state_size(n)
    out = np.zeros((pool.shape[0], sz), np.uint8)
    for i, pre in enumerate(pool):
        diag = anti = 0
        cols = 0
        for r, c in enumerate(pre):
            c = int(c)
            cols |= 1 << c
            diag |= (1 << c) << r
            anti |= (1 << c) << (n - r)
        out[i, 0:8] = np.frombuffer(np.uint64(diag).tobytes(), np.uint8)
        out[i, 8:16] = np.frombuffer(np.uint64(anti).tobytes(), np.uint8)
        out[i, 16:20] = np.frombuffer(np.uint32(cols).tobytes(), np.uint8)
        out[i, 20] = lock
        out[i, 21:21 + n] = 0xFF
        out[i, 21:21 + lock] = pre
    return out


def solver_driver(n, states, locked_row_end, device=0):
    """nqm_solver_driver: the reference's gpu_solver_driver argument layout."""
    L = load()
    states = np.ascontiguousarray(states, dtype=np.uint8)
    cfg = GpuConfig(device, 0)
    tot = ctypes.c_uint64()
    rc = L.nqm_solver_driver(n, states.ctypes.data, states.shape[0], locked_row_end,
                             ctypes.byref(cfg), 1, ctypes.byref(tot))
    if rc != OK:
        raise NqmError(rc, "nqm_solver_driver")
    return tot.value


# ---- subtree-size predictor ---------------------------------------------------

FP = ctypes.POINTER(ctypes.c_float)


class Mlp:
    """nqm_mlp: 12 -> 64 -> 64 -> 1 predictor of log2(1 + subtree nodes)."""

    def __init__(self, ctx, seed=1):
        self.ctx = ctx
        self.lib = ctx.lib
        h = VP()
        ctx._check(self.lib.nqm_mlp_create(ctx.handle, seed, ctypes.byref(h)))
        self.handle = h.value
        self.nparams = int(self.lib.nqm_mlp_param_count())
        self.width = int(self.lib.nqm_mlp_feature_width())

    def close(self):
        if self.handle:
            self.lib.nqm_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(self.nparams, np.float32)
        self.ctx._check(self.lib.nqm_mlp_get_params(self.handle, p.ctypes.data_as(FP)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        self.ctx._check(self.lib.nqm_mlp_set_params(self.handle, p.ctypes.data_as(FP)))

    def predict(self, d_feats, count):
        out = DeviceBuffer(self.ctx, 4 * max(count, 1))
        try:
            self.ctx._check(self.lib.nqm_mlp_predict(self.handle, d_feats.ptr, count, out.ptr))
            return out.to_host(np.float32, count)
        finally:
            out.free()

    def gradient(self, d_feats, d_labels, count):
        g = np.zeros(self.nparams, np.float32)
        loss = ctypes.c_float()
        self.ctx._check(self.lib.nqm_mlp_gradient(self.handle, d_feats.ptr, d_labels.ptr, count,
                                                  g.ctypes.data_as(FP), ctypes.byref(loss)))
        return g, loss.value

    def train(self, d_feats, d_labels, count, epochs=10, batch=4096, lr=1e-3, seed=7, val_frac=0.2):
        tl, vl = ctypes.c_float(), ctypes.c_float()
        self.ctx._check(self.lib.nqm_mlp_train(self.handle, d_feats.ptr, d_labels.ptr, count, epochs,
                                               batch, lr, seed, val_frac, ctypes.byref(tl),
                                               ctypes.byref(vl)))
        return tl.value, vl.value


def _features_pool(ctx, n, depth, d_pool, count):
    """Device feature matrix (count x 16 halves) of a device pool."""
    width = int(ctx.lib.nqm_mlp_feature_width())
    out = DeviceBuffer(ctx, 2 * width * max(count, 1))
    ctx._check(ctx.lib.nqm_mlp_features_pool(ctx.handle, n, depth,
                                             d_pool.ptr if isinstance(d_pool, DeviceBuffer) else d_pool,
                                             count, out.ptr))
    return out


Context.features_pool = _features_pool
Context.mlp = lambda self, seed=1: Mlp(self, seed)


# ---- serving ------------------------------------------------------------------

class Server:
    """nqm_server: continuous batching of count requests on one context."""

    def __init__(self, ctx, **cfg):
        self.ctx = ctx
        self.lib = ctx.lib
        c = ServerCfg()
        for k, v in cfg.items():
            setattr(c, k, v)
        h = VP()
        ctx._check(self.lib.nqm_server_create(ctx.handle, ctypes.byref(c), ctypes.byref(h)))
        self.handle = h.value

    def close(self):
        if self.handle:
            self.lib.nqm_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clock_ms(self):
        return float(self.lib.nqm_server_clock_ms(self.handle))

    def submit(self, reqs, tags):
        """Offer (n, prefix) requests in order; returns how many were admitted."""
        if not reqs:
            return 0
        arr = make_solves(reqs)
        t = np.ascontiguousarray(tags, np.uint64)
        got = ctypes.c_uint32()
        self.ctx._check(self.lib.nqm_server_submit(self.handle, arr, t.ctypes.data_as(U64P), len(reqs),
                                                   ctypes.byref(got)))
        return got.value

    def step(self):
        act = ctypes.c_uint32()
        self.ctx._check(self.lib.nqm_server_step(self.handle, ctypes.byref(act)))
        return act.value

    def poll(self, max_results=4096):
        buf = (Result * max_results)()
        got = ctypes.c_uint32()
        self.ctx._check(self.lib.nqm_server_poll(self.handle, buf, max_results, ctypes.byref(got)))
        return [(r.tag, r.count, r.nodes, r.rounds, r.admit_ms, r.done_ms) for r in buf[:got.value]]

    def info(self):
        a, q, r = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint64()
        self.ctx._check(self.lib.nqm_server_info(self.handle, ctypes.byref(a), ctypes.byref(q),
                                                 ctypes.byref(r)))
        return {"active": a.value, "queued_items": q.value, "rounds": r.value}

    def round_stats(self):
        """Per-round maxima so far (nqm_server_round_stats): rounds, lanes,
        parked_max, handed_back_max, hand_back_room."""
        st = Stats()
        self.ctx._check(self.lib.nqm_server_round_stats(self.handle, ctypes.byref(st)))
        return {k: getattr(st, k) for k in ("rounds", "lanes", "parked_max", "handed_back_max", "hand_back_room")}


Context.server = lambda self, **cfg: Server(self, **cfg)
