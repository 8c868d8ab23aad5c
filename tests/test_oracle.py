"""CPU tests of the oracle: pinned to the golden fixtures written by the compiled reference, to the
compiled reference itself where oracle/_ref is built, and to what the mathematics fixes
(brute force on tiny inputs, closed forms, invariants)."""
import glob
import itertools
import os
import sys

import numpy as np
import pytest

import oracle as O
import rlc
from graphs import cliques, from_pairs, random_weighted, star

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
if GOLD not in sys.path:
    sys.path.insert(0, GOLD)
FIXTURES = sorted(glob.glob(os.path.join(GOLD, "*.npz")))
SYNC_SETTINGS = {  # name -> (slots, pickstep, merge_degree, merge_parts)
    "sync_mg8": (8, 8, 0, 64), "sync_bm": (1, 8, 0, 64), "sync_mg8_p0": (8, 0, 0, 64),
    "sync_mg8_m64": (8, 8, 6, 64), "sync_mg8_m256": (8, 8, 12, 256), "sync_bm_m64": (1, 8, 6, 64),
}


def load_fixture(path):
    z = np.load(path, allow_pickle=False)
    if "offsets" in z:
        g = rlc.Graph(z["offsets"], z["edges"], z["weights"] if "weights" in z else None)
    else:
        kind, n, b, d, f, s = z["generator"].tolist()
        assert kind == "sbm"
        g = rlc.gen_sbm(int(n), int(b), float(d), float(f), int(s))
    return z, g


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_oracle_matches_reference_fixtures(path):
    z, g = load_fixture(path)
    from make_fixtures import csr_digest  # noqa: F401  (same digest as the fixture writer)

    assert csr_digest(g) == str(z["csr_sha256"]), "generator output drifted from the fixture"
    for multi, name in ((True, "lowmem_mg8"), (False, "lowmem_bm")):
        lab, it, _, _ = O.lowmem_omp(g, multi, threads=1)
        assert it == int(z[name + "_iterations"])
        assert np.array_equal(lab, z[name + "_labels"]), name
    lab, it, _, _ = O.rak_omp(g, threads=1)
    assert it == int(z["rak_iterations"]) and np.array_equal(lab, z["rak_labels"])
    unit = g.weights is None or bool(np.all(g.weights == 1.0))
    for name, (slots, pick, md, mp) in SYNC_SETTINGS.items():
        if name + "_labels" not in z:
            continue
        if md > 0 and slots == 8 and not unit:
            # the reference rescan sums in scan order; the merge-tree mode sums in tree order,
            # which agree exactly only when every weight is an integer
            continue
        lab, it, st = O.lpa_sync(g, slots=slots, pickstep=pick, merge_degree=md, merge_parts=mp)
        assert it == int(z[name + "_iterations"]), name
        assert np.array_equal(lab, z[name + "_labels"]), name
        assert np.array_equal(st, z[name + "_stats"]), name
    mg = z["lowmem_mg8_labels"]
    assert O.modularity(g, mg) == float(z["modularity"])
    assert O.communities(mg) == int(z["communities"])


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref (compiled reference) not built here")
@pytest.mark.parametrize("seed", range(4))
def test_oracle_matches_compiled_reference_live(seed):
    r = np.random.default_rng(seed)
    g = rlc.gen_sbm(int(r.integers(500, 4000)), int(r.integers(2, 40)), float(r.uniform(3, 20)),
                    float(r.uniform(0.5, 0.95)), seed)
    gw = random_weighted(800, 4000, seed)
    for gg in (g, gw):
        for multi in (True, False):
            a = O.lowmem_omp(gg, multi, threads=1)
            b = O.ref_lowmem_omp(gg, multi, threads=1)
            assert a[1] == b[1] and np.array_equal(a[0], b[0])
        a, b = O.rak_omp(gg, threads=1), O.ref_rak_omp(gg, threads=1)
        assert a[1] == b[1] and np.array_equal(a[0], b[0])
        for slots, pick in itertools.product((8, 1), (0, 8)):
            a = O.lpa_sync(gg, slots=slots, pickstep=pick)
            b = O.ref_lpa_sync(gg, slots=slots, pickstep=pick)
            assert a[1] == b[1] and np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2])
        lab = O.lowmem_omp(gg, True)[0]
        assert O.modularity(gg, lab) == O.ref_modularity(gg, lab)
        assert O.communities(lab) == O.ref_communities(gg, lab)
    # merge-tree mode on unit weights, where tree-ordered and scan-ordered rescans agree exactly
    for md, mp in ((4, 64), (8, 256), (1, 64)):
        for slots in (8, 1):
            a = O.lpa_sync(g, slots=slots, merge_degree=md, merge_parts=mp)
            b = O.ref_lpa_sync(g, slots=slots, merge_degree=md, merge_parts=mp)
            assert a[1] == b[1] and np.array_equal(a[0], b[0])


# ---- what the mathematics fixes ---------------------------------------------------------------------
def exact_choice(labels, weights):
    """Exact weighted mode, ties to the label whose running total first reached the maximum."""
    tot = {}
    for c, w in zip(labels, weights):
        tot[c] = tot.get(c, 0.0) + float(w)
    best = max(tot.values()) if tot else 0.0
    return {c for c, t in tot.items() if t == best}, best


@pytest.mark.parametrize("seed", range(200))
def test_mg8_equals_exact_mode_when_few_labels(seed):
    # with <= 8 distinct labels no decrement happens: MG8 + rescan is the exact weighted mode
    r = np.random.default_rng(seed)
    k = int(r.integers(1, 9))
    labels = r.choice(np.arange(1000, 1000 + 50), size=k, replace=False)
    seq = r.choice(labels, size=int(r.integers(1, 60)))
    w = r.integers(1, 4, len(seq)).astype(np.float32)  # integer weights: sums are exact
    c, x = O.sketch_choose(seq, w, slots=8)
    best_set, best = exact_choice(seq, w)
    assert c in best_set and x == best
    for parts in (64, 256):  # the merge tree keeps all <= 8 labels too
        c2, x2 = O.sketch_choose(seq, w, slots=8, parts=parts)
        assert c2 in best_set and x2 == best


@pytest.mark.parametrize("seed", range(100))
def test_mg8_keeps_heavy_label(seed):
    # Misra-Gries (clamped) guarantee: a label with more than 1/9 of the total weight survives
    r = np.random.default_rng(seed)
    seq = r.integers(0, 60, 200)
    heavy = 77
    pos = r.choice(200, size=40, replace=False)
    seq[pos] = heavy
    w = np.ones(200, np.float32)
    c, x = O.sketch_choose(seq, w, slots=8)
    assert c == heavy and x == 40.0


@pytest.mark.parametrize("seed", range(100))
def test_bm_returns_strict_majority(seed):
    r = np.random.default_rng(seed)
    n = int(r.integers(1, 80))
    seq = r.integers(0, 10, n)
    w = r.integers(1, 5, n).astype(np.float32)
    tot = {}
    for c, x in zip(seq, w):
        tot[c] = tot.get(c, 0) + x
    c, x = O.sketch_choose(seq, w, slots=1)
    maj = [k for k, t in tot.items() if t * 2 > w.sum()]
    if maj:
        assert c == maj[0]
    assert x > 0


def test_sketch_hand_cases():
    # 9 distinct labels of weight 1: the 9th decrements all eight slots to zero, but the slots
    # keep their labels and the rescan recovers weight 1 for each; labels 0..7 sit in slots
    # 7..0, so the first (lowest) slot holds label 7
    c, x = O.sketch_choose(np.arange(9), np.ones(9, np.float32), slots=8)
    assert (c, x) == (7, 1.0)
    # highest free slot: labels land in slots 7, 6, ...; ties go to the lowest slot = latest first-seen
    c, x = O.sketch_choose(np.array([5, 6, 7]), np.ones(3, np.float32), slots=8)
    assert (c, x) == (7, 1.0)
    # Boyer-Moore: a b b -> (a,1), tie replaces -> (b,1), match -> (b,2)
    c, x = O.sketch_choose(np.array([1, 2, 2]), np.ones(3, np.float32), slots=1)
    assert (c, x) == (2, 2.0)


def brute_modularity(g, lab):
    n = g.n
    W = 0.0
    tot, intra = {}, 0.0
    for u in range(n):
        for i in range(int(g.offsets[u]), int(g.offsets[u + 1])):
            v = int(g.edges[i])
            w = 1.0 if g.weights is None else float(g.weights[i])
            W += w
            tot[lab[u]] = tot.get(lab[u], 0.0) + w
            if lab[u] == lab[v]:
                intra += w
    return intra / W - sum((t / W) ** 2 for t in tot.values())


def test_modularity_closed_forms():
    g = cliques([5, 5], bridges=False)
    lab = np.array([0] * 5 + [5] * 5, np.uint32)
    # two disjoint equal cliques: Q = 1 - 2 * (1/2)^2 = 0.5
    assert abs(O.modularity(g, lab) - 0.5) < 1e-15
    # all in one community: Q = 0
    assert abs(O.modularity(g, np.zeros(10, np.uint32))) < 1e-15
    gw = random_weighted(60, 200, 3)
    for seed in range(5):
        lab = np.random.default_rng(seed).integers(0, 7, gw.n).astype(np.uint32)
        assert abs(O.modularity(gw, lab) - brute_modularity(gw, lab)) < 1e-12
        big = (lab.astype(np.uint64) * 1000003 + 12345).astype(np.uint32)  # labels >= n: same partition
        assert abs(O.modularity(gw, big) - O.modularity(gw, lab)) < 1e-12
    assert O.communities(np.array([3, 3, 9, 1, 9], np.uint32)) == 3


def test_sync_semantics_small():
    # disjoint cliques converge to one label each; isolated vertices keep their own label
    g = cliques([4, 6, 9], bridges=False)
    lab, it, _ = O.lpa_sync(g, pickstep=0)
    assert [len(set(lab[a:b].tolist())) for a, b in ((0, 4), (4, 10), (10, 19))] == [1, 1, 1]
    g2 = from_pairs(4, [(0, 1)])
    lab2, _, st = O.lpa_sync(g2)
    assert lab2[2:].tolist() == [2, 3]
    # a thread count never changes the synchronous result
    g3 = rlc.gen_sbm(3000, 20, 8.0, 0.8, 2)
    a = O.lpa_sync(g3, threads=1)
    b = O.lpa_sync(g3, threads=8)
    assert a[1] == b[1] and np.array_equal(a[0], b[0])
    # star: pick-less first iteration moves every leaf to the hub label 0 (smaller), the hub stays
    gs = star(50)
    lab, it, st = O.lpa_sync(gs)
    assert lab.tolist() == [0] * 51


def test_empty_graph_runs_to_the_cap():
    # 0/0 is not <= tolerance: like the reference loop, an empty graph runs every iteration
    g = rlc.Graph(np.zeros(1, np.uint64), np.zeros(0, np.uint32), None)
    lab, it, st = O.lpa_sync(g, max_iterations=5)
    assert it == 5 and len(lab) == 0


# ---- the exact hashtable vote (slots = 0: rakStaticCuda's choice) -----------------------------------
def exact_sums(labels, weights):
    d = {}
    for c, w in zip(labels, weights):
        d[int(c)] = np.float32(d.get(int(c), np.float32(0)) + np.float32(w))
    return d


@pytest.mark.parametrize("seed", range(30))
def test_exact_vote_picks_a_maximum(seed):
    # the chosen label's weight is the exact sum of its weights (in list order), and no label has more
    r = np.random.default_rng(seed)
    k = int(r.integers(1, 200))
    labels = r.integers(0, int(r.integers(1, 40)), k).astype(np.uint32)
    weights = r.choice([0.5, 1.0, 1.5, 2.0], k).astype(np.float32)
    c, w = O.sketch_choose(labels, weights, slots=0)
    sums = exact_sums(labels, weights)
    assert c in sums and np.float32(w) == sums[c]
    assert w == max(sums.values())


def test_exact_vote_tie_goes_to_first_slot():
    # with H = nextPow2(deg) - 1 slots and keys label+1 whose residues mod H differ (no collision),
    # the first slot of maximum weight wins: the smallest (label + 1) % H among tied labels
    labels = np.array([9, 4, 2, 6, 9, 4, 2, 6], np.uint32)  # deg 8 -> H = 15: residues 10, 5, 3, 7
    c, w = O.sketch_choose(labels, np.ones(8, np.float32), slots=0)
    assert (c, w) == (2, 2.0)
    # deg 6 -> H = T = 7: label 14 -> key 15 -> slot 1; label 0 -> key 1 collides at slot 1, probes
    # i = 1 + 1 -> slot 2; label 5 -> key 6 -> slot 6. All tied at 2: slot 1 (label 14) wins.
    labels = np.array([14, 0, 14, 0, 5, 5], np.uint32)
    c, w = O.sketch_choose(labels, np.ones(6, np.float32), slots=0)
    assert (c, w) == (14, 2.0)


def test_exact_vote_capacity_rule():
    # H = nextPow2(deg) - 1 with nextPow2(x) = 2^bitlength(x): one neighbour gives one slot (found);
    # no neighbour gives no slot (nothing found)
    c, w = O.sketch_choose(np.array([7], np.uint32), np.array([2.0], np.float32), slots=0)
    assert (c, w) == (7, 2.0)
    c, w = O.sketch_choose(np.zeros(0, np.uint32), np.zeros(0, np.float32), slots=0)
    assert c == 0xFFFFFFFF


def test_exact_vote_sync_run_quality():
    # the synchronous exact vote on cliques joined by single edges: one label per clique
    from graphs import cliques
    g = cliques([6, 7, 8, 9])
    lab, it, _ = O.lpa_sync(g, slots=0, pickstep=4)
    starts = np.cumsum([0, 6, 7, 8])
    for s, n in zip(starts, [6, 7, 8, 9]):
        assert len(set(lab[s:s + n].tolist())) == 1
