"""Seeded synthetic workloads: the mixed batch and the serving request stream.

Everything is drawn from splitmix64 so the same seed gives the same solves
on every machine. Pure Python; no device needed.
"""
import math

MASK64 = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014)."""

    def __init__(self, seed):
        self.state = seed & MASK64

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def below(self, m):
        """Integer in [0, m) (modulo draw; the bias is < m / 2^64)."""
        return self.next() % m

    def uniform(self):
        """Float in (0, 1) from the top 53 bits."""
        return ((self.next() >> 11) + 0.5) / float(1 << 53)


def valid_prefixes(n, k):
    """All valid placements of k queens in rows 0..k-1 of an n x n board, in
    lexicographic order (a queen per row, no shared column or diagonal)."""
    out = []
    cols = []

    def rec(row, used, ld, rd):
        if row == k:
            out.append(tuple(cols))
            return
        free = ~(used | ld | rd) & ((1 << n) - 1)
        c = 0
        while free:
            if free & 1:
                b = 1 << c
                cols.append(c)
                rec(row + 1, used | b, ((ld | b) << 1) & ((1 << n) - 1), (rd | b) >> 1)
                cols.pop()
            free >>= 1
            c += 1

    rec(0, 0, 0, 0)
    return out


_prefix_cache = {}


def random_prefix(rng, n, k):
    key = (n, k)
    if key not in _prefix_cache:
        _prefix_cache[key] = valid_prefixes(n, k)
    pool = _prefix_cache[key]
    return pool[rng.below(len(pool))] if pool else ()


def mixed_batch(seed=0x5EED0512, count=512, n_lo=8, n_hi=17, k_max=4):
    """The batched workload: `count` solves, n ~ U{n_lo..n_hi}, k ~ U{0..k_max}
    pre-placed queens drawn uniformly from the valid k-prefixes of that n."""
    rng = SplitMix64(seed)
    solves = []
    for _ in range(count):
        n = n_lo + rng.below(n_hi - n_lo + 1)
        k = rng.below(k_max + 1)
        solves.append((n, tuple(random_prefix(rng, n, k))))
    return solves


def serving_stream(seed=0x5E4F1FE5, count=2000, n_lo=10, n_hi=20, prefix_frac=0.30,
                   k_lo=1, k_hi=3, rate=1.0):
    """Serving requests: (arrival_time_s, n, prefix). Poisson arrivals at
    `rate` requests/s; n from n_lo..n_hi with weight proportional to 1/n^2;
    a fraction `prefix_frac` carries a random valid prefix of k_lo..k_hi."""
    rng = SplitMix64(seed)
    ns = list(range(n_lo, n_hi + 1))
    w = [1.0 / (n * n) for n in ns]
    tot = sum(w)
    cdf = []
    acc = 0.0
    for x in w:
        acc += x / tot
        cdf.append(acc)
    t = 0.0
    out = []
    for _ in range(count):
        t += -math.log(rng.uniform()) / rate
        u = rng.uniform()
        n = ns[-1]
        for v, c in zip(ns, cdf):
            if u <= c:
                n = v
                break
        prefix = ()
        if rng.uniform() < prefix_frac:
            k = k_lo + rng.below(k_hi - k_lo + 1)
            prefix = tuple(random_prefix(rng, n, k))
        out.append((t, n, prefix))
    return out
