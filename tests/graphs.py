"""Seeded test graphs shared by the CPU and GPU test modules."""
import numpy as np

import tcamd


def er(n=2000, m=20000, seed=1):
    return tcamd.gen_er(n, m, seed)


def rmat(scale=12, ef=16, seed=22):
    return tcamd.gen_rmat(scale, ef << scale, 0.57, 0.19, 0.19, seed)


def ba(n=3000, k=8, seed=3):
    return tcamd.gen_ba(n, k, seed)


def rgg(n=3000, deg=20, seed=4):
    return tcamd.gen_rgg(n, tcamd.rgg_radius(n, deg), seed)


def hub(n=6000, m=60000, seed=7):
    """ER on [0, n) plus vertex n joined to all of it: under ID orientation
    vertex n has out-degree n (a hub longer than any LDS slice)."""
    s, d = tcamd.gen_er(n, m, seed)
    hs = np.full(n, n, np.uint32)
    hd = np.arange(n, dtype=np.uint32)
    return np.concatenate([s, hs]), np.concatenate([d, hd])


def clique(k):
    iu = np.triu_indices(k, 1)
    return iu[0].astype(np.uint32), iu[1].astype(np.uint32)


def messy():
    """Duplicates in both directions, self-loops, isolated ids."""
    s = np.array([0, 1, 1, 2, 2, 0, 3, 3, 5, 5, 9, 9, 2, 4, 4, 1], np.uint32)
    d = np.array([1, 0, 2, 1, 0, 2, 3, 0, 4, 4, 9, 2, 9, 9, 2, 9], np.uint32)
    return s, d


CASES = {
    "er": er,
    "rmat12": rmat,
    "ba": ba,
    "rgg": rgg,
    "hub": hub,
    "k6": lambda: clique(6),
    "messy": messy,
}
