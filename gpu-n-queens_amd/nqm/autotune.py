"""Autotuner front end: sweeps expansion depth x waves per CU x LDS stack rows
through nqm_autotune (C++) and caches the winner per device and workload in a
JSON file, so later runs start from the tuned configuration."""
import ctypes
import json
import os

from . import Tune, TuneResult, TuneSpace, make_solves

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune_cache.json")


def _fill(arr, values):
    for i, v in enumerate(list(values)[:8]):
        arr[i] = v


def sweep(ctx, solves, depths=(3, 4, 5, 6, 7), waves=(4, 8, 12, 16), stack_rows=(), repeats=3,
          base_flags=0):
    """Returns (best Tune, [(depth, waves_per_cu, stack_rows, median_ms), ...])."""
    sp = TuneSpace()
    _fill(sp.depths, depths)
    _fill(sp.waves_per_cu, waves)
    _fill(sp.stack_rows, stack_rows)
    sp.repeats = repeats
    sp.base_flags = base_flags
    arr = make_solves(solves)
    best = Tune()
    res = (TuneResult * 512)()
    nres = ctypes.c_uint32()
    ctx._check(ctx.lib.nqm_autotune(ctx.handle, arr, len(solves), ctypes.byref(sp), ctypes.byref(best),
                                    res, 512, ctypes.byref(nres)))
    rows = [(r.depth, r.waves_per_cu, r.stack_rows, r.ms) for r in res[:nres.value]]
    return best, rows


def key(ctx, workload):
    info = ctx.device_info()
    return f"cus{info['cus']}-clk{info['clock_khz']}-{workload}"


def load_cache(path=CACHE):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def store(ctx, workload, best, rows, path=CACHE):
    c = load_cache(path)
    c[key(ctx, workload)] = {"depth": best.depth, "waves_per_cu": best.waves_per_cu,
                             "stack_rows": best.stack_rows, "trials": rows}
    with open(path, "w") as f:
        json.dump(c, f, indent=1)


def cached_tune(ctx, workload, path=CACHE):
    e = load_cache(path).get(key(ctx, workload))
    if not e:
        return None
    t = Tune()
    t.depth, t.waves_per_cu, t.stack_rows = e["depth"], e["waves_per_cu"], e["stack_rows"]
    return t
