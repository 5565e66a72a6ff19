"""Synthetic workloads of the batch and serving forms, drawn from pinned seeds.

uneven_batch(): instances (n, d, M) of different size for the one-launch work queue:
n = round(exp(U(ln 64, ln 4096))), d = 2 + floor(3 U), M = ceil(alpha n) with alpha = U(1.05, 3),
drawn per instance in that order from numpy's PCG64 generator.

serving_queries(): nbad(n, M) requests for tables of d = 3 up to n_max rows: n uniform in
[n_min, n_max], M uniform in [n + 1, 4 n], and exponential inter-arrival gaps of mean 1/rate.
"""
import math

import numpy as np


def uneven_batch(seed=0, count=256, n_lo=64, n_hi=4096):
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        n = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))))
        d = 2 + int(math.floor(3 * rng.uniform()))
        alpha = rng.uniform(1.05, 3.0)
        out.append((n, d, int(math.ceil(alpha * n))))
    return out


def serving_queries(seed=1, count=100000, n_min=32, n_max=2048, rate=1000.0):
    """Returns (n, M, arrival_seconds) arrays."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = rng.integers(n_min, n_max + 1, size=count)
    M = n + 1 + np.floor(rng.uniform(size=count) * (3 * n)).astype(np.int64)
    t = np.cumsum(rng.exponential(1.0 / rate, size=count))
    return n.astype(np.int64), M, t
