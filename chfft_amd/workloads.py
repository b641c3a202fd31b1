"""Synthetic workloads of the benchmark configurations (seeded, host-side)."""
import numpy as np


def smooth235_upto(limit):
    """All 2^a 3^b 5^c <= limit, sorted."""
    out = []
    p2 = 1
    while p2 <= limit:
        p3 = p2
        while p3 <= limit:
            p5 = p3
            while p5 <= limit:
                out.append(p5)
                p5 *= 5
            p3 *= 3
        p2 *= 2
    return np.array(sorted(out), dtype=np.int64)


def primes_upto(limit):
    sieve = np.ones(limit + 1, dtype=bool)
    sieve[:2] = False
    for i in range(2, int(limit ** 0.5) + 1):
        if sieve[i]:
            sieve[i * i::i] = False
    return np.nonzero(sieve)[0].astype(np.int64)


def uneven_lengths(count=10000, lo=16, hi=1 << 20, cap=1 << 28, seed=0x5EED0005):
    """Lengths of the uneven-batch workload: log-uniform draws on [lo, hi],
    alternately rounded up to the next 2*3*5-smooth number and to the next
    prime (<= the largest prime <= hi). The total is capped at `cap` points
    (the uncapped expectation is ~3.5x larger) by a running budget: entry i
    may use at most cap * (i + 1) / count - (points so far); a draw over
    budget is redrawn (up to 64 draws per entry), and if none fits the entry
    becomes the largest length of its kind within budget (at least the
    smallest). Large lengths stay possible whenever small draws have left
    budget. The list is a deterministic function of the seed."""
    rng = np.random.default_rng(seed)
    sm = smooth235_upto(hi)
    sm = sm[sm >= lo]
    pr = primes_upto(hi)
    pr = pr[pr >= lo]
    draws = np.exp(rng.uniform(np.log(lo), np.log(hi), size=(count, 64)))
    out, total = [], 0
    for i in range(count):
        table = sm if i % 2 == 0 else pr
        cand = table[np.minimum(np.searchsorted(table, draws[i]), len(table) - 1)]
        room = cap * (i + 1) // count - total
        ok = np.nonzero(cand <= room)[0]
        if len(ok):
            n = int(cand[ok[0]])
        else:
            k = max(int(np.searchsorted(table, room, side="right")) - 1, 0)
            n = int(table[k])
        out.append(n)
        total += n
    return out


def _smooth7(n):
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


UNEVEN_CLASSES = ("all", "fused", "smooth4", "blue4_small", "blue4_big")


def uneven_class(lens, name):
    """The lengths of one size class of an uneven batch, by how the FP32
    multi-length plan runs them: "fused" (7-smooth up to 16384 points, other
    lengths up to 8192: one workgroup each), "smooth4" (7-smooth above 16384:
    four-step), "blue4_small" / "blue4_big" (four-step Bluestein, N up to /
    above 65536), or "all"."""
    if name == "all":
        return list(lens)
    sel = {
        "fused": lambda n: n <= 16384 if _smooth7(n) else n <= 8192,
        "smooth4": lambda n: _smooth7(n) and n > 16384,
        "blue4_small": lambda n: not _smooth7(n) and 8192 < n <= 65536,
        "blue4_big": lambda n: not _smooth7(n) and n > 65536,
    }[name]
    return [n for n in lens if sel(n)]
