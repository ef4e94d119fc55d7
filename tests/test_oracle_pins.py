"""Known-answer tests that pin three oracle functions the generic checks leave
open, each derived by hand from the reference source, and a check that each
of them catches a plausible mistake: the mistake is planted in a copy of
oracle.cpp (on the CPU, in a temporary directory), the copy is built, and the
known-answer test must fail against it while passing against the unchanged
source.

* A-direction's outer threshold step, edge_directing/A-direction-seri.cpp:77,127:
  the threshold grows by the integer part of threshold0, not by threshold0.
* Fox's lanes-per-edge bin boundaries, Fox/tc.cu:34-63,316-331.
* Hu's degree<=1 pruning and id compaction, Hu/preTC.cu:135-169.
* Bisson's A-direction relabel, Bisson/A-direction.cpp:34-163,287-297: the
  degree record a cascade push writes into the frontier index's slot.
* The OpenMP A-direction program's schedule, edge_directing/A-direction.cpp:
  38,96,139: threshold0, the 200-vertex cascade gate, the six-level cap and
  the thre_param step.
"""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
W = 0xFFFFFFFF  # a wiped endpoint (Hu/preTC.cu:151-160 writes -1)


def kat_adir_threshold_step():
    """K4 on ids 0-3 plus a self-loop record on id 12.

    nodeNum is the largest id + 1 over all records, the self-loop included
    (A-direction-seri.cpp:187-193), so n = 13 and ids 4..12 are isolated;
    the self-loop itself is dropped (:209-210).  totalEdge = 2 * 6 = 12, so
    threshold0 = 12/13 + 1 = 1.923 and its integer part is 1 (:37,:77).
    Doubled degrees (:44-47): 6 on the K4, 0 elsewhere.

    round 0, threshold 1.923: ids 4..12 get label 0; their cascade is empty
                              and uses label 1.
    rounds 1-4, thresholds 2.923, 3.923, 4.923, 5.923: nothing <= them;
                              labels 2, 3, 4, 5.
    round 5, threshold 6.923: the K4 (6) gets label 6.

    A float step (threshold += threshold0) would reach 7.69 >= 6 in round 3
    and give the K4 label 4."""
    s = np.array([0, 0, 0, 1, 1, 2, 12], np.uint32)
    d = np.array([1, 2, 3, 2, 3, 3, 12], np.uint32)
    n = orc.num_nodes(s, d)
    assert n == 13
    srp, scol = orc.build_sym(s, d, n)
    assert orc.adir_priority(n, srp, scol).tolist() == [6, 6, 6, 6] + [0] * 9


def kat_fox_bins():
    """Sorted per-edge work with binMaxWork = 1000 (Fox/tc.cu:316-331).

    maxPowerOf8 = ln 1000 / ln 8 = 3.32193.  For i = 2, 4, 8, 16, 32:
    index = 2.32193 (i - 1) / 32 + 1 = 1.07256, 1.21768, 1.50792, 2.08840,
    3.24937, and partitionPoint = (int) 8^index = 9, 12, 23, 76, 859.
    binarySearchValue(.., -1) (:34-63) returns the first position holding
    the value (9 -> 2, 12 -> 5, 23 -> 7, 76 -> 10, 859 -> 13), or the
    insertion point when it is absent.  binStartPos[i] and
    binStartPos[i/2 + 1] both take it; binStartPos[1] = 0 and
    binStartPos[33] = size = 16."""
    w = np.array([1, 5, 9, 9, 10, 12, 20, 23, 23, 50, 76, 77, 500, 859, 860, 1000], np.int32)
    b = orc.fox_bins(w, 32).tolist()
    want = {1: 0, 2: 2, 3: 5, 4: 5, 5: 7, 8: 7, 9: 10, 16: 10, 17: 13, 32: 13, 33: 16}
    assert {i: b[i] for i in want} == want
    # the bins of 1, 2, 4, 8, 16, 32 lanes per edge (:334-336)
    assert [b[i + 1] - b[i] for i in (1, 2, 4, 8, 16, 32)] == [2, 3, 2, 3, 3, 3]


def kat_prune_numbering():
    """Hu/preTC.cu:135-169 on 9 raw records over ids 0..9.

    Raw degrees (duplicates counted): 0:1 1:2 2:3 3:4 4:0 5:3 6:2 7:1 8:1
    9:1, so 0, 4, 7, 8, 9 are removed and toBeMini = [1,1,1,1,2,2,2,3,4,5]
    (inclusive count).  A kept id x > 0 becomes x - toBeMini[x-1]: 1->0,
    2->1, 3->2, 5->3, 6->4.  The source is tested first: a record whose
    source is removed gets src = -1 and keeps its destination as it was
    (8-9 -> (-1, 9)); one whose destination alone is removed gets dst = -1
    and keeps its source id unchanged (5-7 -> (5, -1)): the loop continues
    before renumbering either.  nodeNum becomes 10 - 5 = 5."""
    s = np.array([0, 1, 2, 3, 3, 5, 6, 5, 8], np.uint32)
    d = np.array([2, 2, 3, 1, 5, 6, 3, 7, 9], np.uint32)
    deg = orc.degree_raw(s, d, 10)
    assert deg.tolist() == [1, 2, 3, 4, 0, 3, 2, 1, 1, 1]
    ps, pd, n2 = orc.prune_deg1(s, d, 10, deg)
    assert n2 == 5
    assert ps.tolist() == [W, 0, 1, 2, 2, 3, 4, 5, W]
    assert pd.tolist() == [2, 1, 2, 0, 3, 4, 2, W, 9]
    # id 0 kept: it stays 0 (the `src > 0` guard)
    s = np.array([0, 1, 2, 3], np.uint32)
    d = np.array([1, 2, 0, 4], np.uint32)
    ps, pd, n2 = orc.prune_deg1(s, d, 5, orc.degree_raw(s, d, 5))
    assert (ps.tolist(), pd.tolist(), n2) == ([0, 1, 2, W], [1, 2, 0, 4], 3)


def kat_bisson_relabel():
    """Bisson/A-direction.cpp on two hand-traced record sets.

    Records (4,4), (3,1), (4,2), (2,3) over ids 0..4: the self-loop is
    dropped (:227-228), leaving the path 1-3-2-4 and isolated 0.  Doubled
    degrees [0,2,4,4,2], totalEdge 6, threshold 6/5 + 1 = 2.2.
    Scan: 0, 1, 4 get priority 0, records 0, 2, 2 (:91-98).
    Cascade, label 1 (:101-124): frontier index 1 (vertex 1) takes vertex 3
    from 3 to 2 on its second allEdges entry and pushes it, writing 2 into
    degreeRecord[1]; index 2 (vertex 4) pushes vertex 2 and writes
    degreeRecord[2] = 2.  Vertex 3's own record stays 0 (:263 memset).
    Order by (priority, record, id) (:34-42): 0, 1, 4, 3, 2, so lookUpTable
    = 4 - rank: [4, 3, 0, 1, 2].  Recording at the pushed vertex instead
    would order 2 before 3 and give [4, 3, 1, 0, 2].

    K4 on 0-3, path 4-5-6 and a self-loop on 7: threshold 16/8 + 1 = 3;
    scan peels 4, 6, 7 (records 2, 2, 0), vertex 4 pushes 5 (record slot 0,
    later overwritten by the K4's scan record 6), threshold 6 peels the K4
    with label 3: order 7, 4, 6, 5, 0, 1, 2, 3 -> [3, 2, 1, 0, 6, 4, 5, 7]."""
    s = np.array([4, 3, 4, 2], np.uint32)
    d = np.array([4, 1, 2, 3], np.uint32)
    assert orc.relabel(s, d, "adir").tolist() == [4, 3, 0, 1, 2]
    s = np.array([0, 0, 0, 1, 1, 2, 4, 5, 7], np.uint32)
    d = np.array([1, 2, 3, 2, 3, 3, 5, 6, 7], np.uint32)
    assert orc.relabel(s, d, "adir").tolist() == [3, 2, 1, 0, 6, 4, 5, 7]


def _star(k):
    return np.zeros(k, np.uint32), np.arange(1, k + 1, dtype=np.uint32)


def _layered():
    """16 layers of 210 vertices; vertex j of layer i joins j, j+1, j+2 (mod
    210) of layer i+1, and layers 7-8 also j+3."""
    src, dst = [], []
    for i in range(15):
        for j in range(210):
            for t in range(3 + (i == 7)):
                src.append(i * 210 + j)
                dst.append((i + 1) * 210 + (j + t) % 210)
    return np.array(src, np.uint32), np.array(dst, np.uint32)


def _omp_prio(s, d, thre_param):
    n = orc.num_nodes(s, d)
    srp, scol = orc.build_sym(s, d, n)
    return orc.adir_omp_priority(n, srp, scol, thre_param)


def kat_adir_omp_schedule():
    """edge_directing/A-direction.cpp:36-161 on three hand-traced graphs.

    Star with 150 leaves, thre_param 10: totalEdge 300, threshold0 =
    2 * 300 / 151 = 3.97 (:38), threshold_old 3.  Round 0 peels the leaves
    (doubled degree 2); 150 is not above 200 (:96), so no cascade and the
    centre keeps 300.  Each round adds 3 * 10 / 5 = 6 (:139) and uses one
    label: 3.97 + 6k >= 300 first at k = 50, the centre's label.
    Star with 300 leaves, thre_param 5: threshold0 = 1200 / 301 = 3.99; the
    300 leaves cascade, the centre's 600 falls to 3 and it takes label 1.
    Layered graph (_layered), thre_param 1: 9660 edges, n = 3360, threshold0
    = 2 * 19320 / 3360 = 11.5, doubled degrees 6 (layers 0, 15), 12, and 14
    (layers 7, 8).  Round 0 peels layers 0 and 15; six cascade levels (the
    cap, :96) peel layers k and 15 - k with label k; layers 6 and 9 are not
    expanded, so 7 and 8 keep 14.  Round 1 (threshold 11.5 + 2.2 = 13.7,
    label 7) peels nothing; round 2 (15.9) gives them label 8."""
    p = _omp_prio(*_star(150), 10)
    assert p[0] == 50 and set(p[1:].tolist()) == {0}
    p = _omp_prio(*_star(300), 5)
    assert p[0] == 1 and set(p[1:].tolist()) == {0}
    p = _omp_prio(*_layered(), 1).reshape(16, 210)
    want = [0, 1, 2, 3, 4, 5, 6, 8, 8, 6, 5, 4, 3, 2, 1, 0]
    assert [sorted(set(r.tolist())) for r in p] == [[w] for w in want]


def test_adir_threshold_step_known_answer():
    kat_adir_threshold_step()


def test_fox_bins_known_answer():
    kat_fox_bins()


def test_prune_numbering_known_answer():
    kat_prune_numbering()


def test_bisson_relabel_known_answer():
    kat_bisson_relabel()


def test_adir_omp_schedule_known_answer():
    kat_adir_omp_schedule()


# (known-answer test, text in oracle.cpp, planted mistake)
MUTANTS = [
    ("kat_adir_threshold_step", "threshold += threshold_old;",
     "threshold += ((float)edgeNum) / nodeNum + 1;"),
    ("kat_adir_threshold_step", ": ((float)edgeNum) / nodeNum + 1;",
     ": ((float)edgeNum) / nodeNum + 2;"),
    ("kat_fox_bins", "int workPerThreadB = 8;", "int workPerThreadB = 4;"),
    ("kat_fox_bins", "(double)(maxPowerOf8 - 1) * (i - 1) / (double)curBinB + 1",
     "(double)(maxPowerOf8 - 1) * i / (double)curBinB + 1"),
    ("kat_fox_bins", "fox_binary_search_value(w, partitionPoint, size, -1)",
     "fox_binary_search_value(w, partitionPoint, size, 1)"),
    ("kat_prune_numbering",
     "if (deg[s] <= 1) { src[i] = kWiped; continue; }\n    if (deg[d] <= 1) { dst[i] = kWiped; continue; }",
     "if (deg[d] <= 1) { dst[i] = kWiped; continue; }\n    if (deg[s] <= 1) { src[i] = kWiped; continue; }"),
    ("kat_prune_numbering", "if (s > 0) src[i] = s - mini[s - 1];",
     "if (s > 0) src[i] = s - mini[s - 1] - 1;"),
    ("kat_prune_numbering", "if (deg[i] <= 1) total++;", "if (deg[i] < 1) total++;"),
    ("kat_prune_numbering", "if (deg[d] <= 1) { dst[i] = kWiped; continue; }",
     "if (deg[d] <= 1) { dst[i] = kWiped; if (s > 0) src[i] = s - mini[s - 1]; continue; }"),
    ("kat_bisson_relabel", "if (rec) rec[f] = rtd[node];", "if (rec) rec[node] = rtd[node];"),
    ("kat_bisson_relabel", "priority == b.priority && degree < b.degree",
     "priority == b.priority && degree > b.degree"),
    ("kat_bisson_relabel", "perm[units[i].id] = n - 1 - i;", "perm[units[i].id] = i;"),
    ("kat_adir_omp_schedule", "omp ? 2 * ((float)edgeNum) / nodeNum :",
     "omp ? ((float)edgeNum) / nodeNum + 1 :"),
    ("kat_adir_omp_schedule", "if (omp) threshold += (double)threshold_old * thre_param / 5;",
     "if (omp) threshold += threshold_old;"),
    ("kat_adir_omp_schedule", "(int)frontier.size() > 200 &&", "(int)frontier.size() > 100 &&"),
    ("kat_adir_omp_schedule", "innerLoop <= 5", "innerLoop <= 6"),
]


def _build(tmp, name, source):
    d = os.path.join(tmp, name)
    os.makedirs(d, exist_ok=True)
    shutil.copy(os.path.join(ROOT, "oracle", "oracle.h"), d)
    with open(os.path.join(d, "oracle.cpp"), "w") as f:
        f.write(source)
    lib = os.path.join(d, "liboracle.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-o", lib,
                    os.path.join(d, "oracle.cpp")], check=True, capture_output=True)
    return lib


def _run_kat(lib, kat):
    code = ("import sys; sys.path[:0] = [sys.argv[1], sys.argv[2]]; "
            "import test_oracle_pins as t; getattr(t, sys.argv[3])()")
    env = dict(os.environ, ORACLE_LIB=lib)
    return subprocess.run([sys.executable, "-c", code, os.path.join(ROOT, "tests"),
                           os.path.join(ROOT, "oracle"), kat], env=env, capture_output=True,
                          text=True, timeout=60).returncode


def test_known_answers_catch_planted_mistakes(tmp_path):
    with open(os.path.join(ROOT, "oracle", "oracle.cpp")) as f:
        src = f.read()
    control = _build(str(tmp_path), "control", src)
    for kat in sorted({m[0] for m in MUTANTS}):
        assert _run_kat(control, kat) == 0, kat
    for i, (kat, text, planted) in enumerate(MUTANTS):
        assert src.count(text) == 1, text
        lib = _build(str(tmp_path), f"m{i}", src.replace(text, planted))
        assert _run_kat(lib, kat) != 0, f"{kat} passes with the planted mistake {planted!r}"
