"""Real-time replay of a request stream against nqm_server (continuous
batching): requests are offered when their arrival time has passed on the
server clock; p50/p99 latency is completion minus arrival."""
import time

import numpy as np


def replay(srv, requests, on_result=None, time_scale=1.0):
    """requests: list of (arrival_s, n, prefix) sorted by arrival. Returns a
    dict tag -> (count, nodes, rounds, latency_ms, admit_ms, done_ms) and the
    run's wall time in ms."""
    pending = list(enumerate(requests))
    i = 0
    out = {}
    t_start = srv.clock_ms()
    waiting = []  # arrived, not yet admitted
    stalled = 0
    while i < len(pending) or waiting or srv.info()["active"]:
        now = srv.clock_ms() - t_start
        while i < len(pending) and pending[i][1][0] * 1e3 * time_scale <= now:
            waiting.append(pending[i])
            i += 1
        if waiting:
            took = srv.submit([(n, tuple(p)) for _, (_, n, p) in waiting], [tag for tag, _ in waiting])
            waiting = waiting[took:]
        info = srv.info()
        # an idle server that refuses a waiting request would spin forever
        stalled = stalled + 1 if (waiting and info["active"] == 0 and info["queued_items"] == 0) else 0
        if stalled > 3:
            raise RuntimeError("server admits nothing while idle")
        if info["active"] == 0 and not waiting:
            if i < len(pending):  # idle: sleep until the next arrival
                nxt = pending[i][1][0] * 1e3 * time_scale - (srv.clock_ms() - t_start)
                if nxt > 0:
                    time.sleep(nxt / 1e3)
            continue
        srv.step()
        for tag, count, nodes, rounds, admit_ms, done_ms in srv.poll():
            arrival_ms = requests[tag][0] * 1e3 * time_scale + t_start
            rec = (count, nodes, rounds, done_ms - arrival_ms, admit_ms, done_ms)
            out[tag] = rec
            if on_result:
                on_result(tag, rec)
    wall = srv.clock_ms() - t_start
    return out, wall


def latency_summary(results, wall_ms):
    lat = np.array([r[3] for r in results.values()], np.float64)
    return {"requests": int(lat.size), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
            "max_ms": float(lat.max()), "wall_s": wall_ms / 1e3,
            "solves_per_s": lat.size / (wall_ms / 1e3)}
