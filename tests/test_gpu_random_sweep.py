"""Seeded random sweep: many small graphs of varied size, density and skew
(uniform, skewed power-law-like, dense blocks, star-heavy, raw records with
duplicates and self-loops) through every orientation, every triangle
kernel, the learned schedule and every k-clique path, each checked against
the oracle bit for bit.  The exact CSR of every orientation is compared too."""
import numpy as np
import pytest

import oracle as orc
import tcamd

pytestmark = pytest.mark.gpu

ORIENTS = {
    "id": tcamd.ORIENT_ID, "degree": tcamd.ORIENT_DEGREE, "tricore": tcamd.ORIENT_DEGREE_TRICORE,
    "adir": tcamd.ORIENT_ADIR, "kcore": tcamd.ORIENT_KCORE,
    "degeneracy": tcamd.ORIENT_DEGENERACY,
}
KERNELS = ["hu", "bisson", "fox", "tricore", "auto"]


def random_graph(seed):
    """Raw records (duplicates, both directions and self-loops allowed)."""
    rng = np.random.default_rng(1000 + seed)
    kind = seed % 5
    n = int(rng.integers(2, 2500))
    if kind == 0:    # uniform
        m = int(rng.integers(1, 12 * n))
        s, d = rng.integers(0, n, m), rng.integers(0, n, m)
    elif kind == 1:  # skewed ids (a few hubs)
        m = int(rng.integers(1, 12 * n))
        s = (n * rng.random(m) ** 3).astype(np.int64)
        d = (n * rng.random(m) ** 2).astype(np.int64)
    elif kind == 2:  # dense blocks plus noise
        b = int(rng.integers(3, 90))
        blocks = int(rng.integers(1, 6))
        ss, dd = [], []
        for i in range(blocks):
            off = int(rng.integers(0, max(1, n - b)))
            iu = np.triu_indices(b, 1)
            keep = rng.random(len(iu[0])) < rng.uniform(0.5, 1.0)
            ss.append(iu[0][keep] + off)
            dd.append(iu[1][keep] + off)
        m = int(rng.integers(0, 3 * n))
        ss.append(rng.integers(0, n, m))
        dd.append(rng.integers(0, n, m))
        s, d = np.concatenate(ss), np.concatenate(dd)
    elif kind == 3:  # stars around a few centres
        c = int(rng.integers(1, 8))
        m = int(rng.integers(1, 6 * n))
        s = rng.integers(0, c, m)
        d = rng.integers(0, n, m)
        m2 = int(rng.integers(0, 2 * n))
        s = np.concatenate([s, rng.integers(0, n, m2)])
        d = np.concatenate([d, rng.integers(0, n, m2)])
    else:            # duplicates and self-loops on purpose
        m = int(rng.integers(1, 6 * n))
        s, d = rng.integers(0, n, m), rng.integers(0, n, m)
        s = np.concatenate([s, d[: m // 3], s[: m // 5]])
        d = np.concatenate([d, s[: m // 3], s[: m // 5]])
    return s.astype(np.uint32), d.astype(np.uint32)


def oracle_csr(s, d, mode):
    n = orc.num_nodes(s, d)
    if mode in ("id", "degree", "tricore"):
        deg = orc.degree_raw(s, d, n)
        m = {"id": orc.ORIENT_ID, "degree": orc.ORIENT_DEGREE,
             "tricore": orc.ORIENT_DEG_TRICORE}[mode]
        return orc.build_csr(*orc.orient_simple(s, d, deg, m), n)
    srp, scol = orc.build_sym(s, d, n)
    if mode == "adir":
        key = orc.adir_priority(n, srp, scol)
        return orc.build_csr(*orc.adir_orient(n, srp, scol, key), n)
    key = orc.kcore(n, srp, scol) if mode == "kcore" else orc.degeneracy(n, srp, scol)[0]
    return orc.build_csr(*orc.kcore_orient(n, srp, scol, key), n)


@pytest.mark.parametrize("seed", range(30))
def test_random_graph_every_path(ctx, seed):
    s, d = random_graph(seed)
    print(f"seed={seed} records={len(s)}", flush=True)
    expect = None
    model = tcamd.Model(ctx, seed)
    for oname, omode in ORIENTS.items():
        rp_o, col_o = oracle_csr(s, d, oname)
        if expect is None:
            expect = orc.count(rp_o, col_o)
        g = ctx.graph(s, d).orient(omode)
        rp, col = g.csr()
        np.testing.assert_array_equal(rp, rp_o, err_msg=f"rowptr {oname}")
        np.testing.assert_array_equal(col, col_o, err_msg=f"col {oname}")
        for k in KERNELS:
            got = g.count(k).count
            assert got == expect, f"seed {seed} {oname}/{k}: {got} != {expect}"
        assert g.count_learned(model).count == expect
        if oname == "degeneracy":
            for kk in (4, 5):
                want = orc.kclique(rp_o, col_o, kk)
                for v in (tcamd.KCL_AUTO, tcamd.KCL_PIVOT, tcamd.KCL_BITMAP, tcamd.KCL_MERGE):
                    assert g.kclique(kk, v).count == want, f"seed {seed} k={kk} variant {v}"
                assert g.kclique_learned(kk, model).count == want
        g.close()
