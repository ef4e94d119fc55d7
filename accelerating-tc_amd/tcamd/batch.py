"""Synthetic batches of many small graphs for batched counting.

A batch of `count` graphs: vertex counts log-uniform in [n_min, n_max],
average degree uniform in [deg_min, deg_max], graph i of kind i % 3
(0 = Erdos-Renyi, 1 = R-MAT (0.57, 0.19, 0.19, 0.05), 2 = random geometric in
the unit square with the radius set for the target degree).  Every graph is
drawn from its own seed, so a batch is reproducible from (count, seed, ...).
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import tcamd

KINDS = ("er", "rmat", "rgg")


def graph_params(count=4096, seed=5, n_min=64, n_max=65536, deg_min=4.0, deg_max=64.0):
    """(kind, n, target average degree, seed) of every graph of the batch."""
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(n_min), np.log(n_max), count))
    n = np.clip(np.rint(n), n_min, n_max).astype(np.int64)
    deg = rng.uniform(deg_min, deg_max, count)
    seeds = rng.integers(1, 2 ** 62, count)
    return [(KINDS[i % 3], int(n[i]), float(deg[i]), int(seeds[i])) for i in range(count)]


def make_one(kind, n, deg, seed):
    """Raw records (src, dst) of one graph with local ids < n."""
    if kind == "er":
        m = int(round(n * deg / 2))
        m = max(1, min(m, n * (n - 1) // 4))  # at most half the possible pairs
        return tcamd.gen_er(n, m, seed)
    if kind == "rmat":
        scale = max(1, int(np.ceil(np.log2(n))))
        want = int(round(n * deg / 2 * (1 << scale) / n))
        s, d = tcamd.gen_rmat(scale, max(1, want), 0.57, 0.19, 0.19, seed)
        keep = (s < n) & (d < n)
        return s[keep], d[keep]
    if kind == "rgg":
        return tcamd.gen_rgg(n, tcamd.rgg_radius(n, min(deg, n - 1.0)), seed)
    raise ValueError(kind)


def make_batch(count=4096, seed=5, threads=None, **kw):
    """List of (src, dst, n) for the batch (generators run on host threads)."""
    params = graph_params(count, seed, **kw)
    threads = threads or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(threads) as ex:
        edges = list(ex.map(lambda p: make_one(*p), params))
    return [(s, d, p[1]) for (s, d), p in zip(edges, params)]
