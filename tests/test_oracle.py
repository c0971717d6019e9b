"""Pins the CPU oracle (oracle/) before anything is checked against it.

Each test compares the oracle with something that does not share its code:
the C++ standard's known answers, the reference's own fixtures and KATs,
closed forms, or a textbook brute force (tests/bruteforce.py).
"""
import os

import numpy as np
import pytest

import oracle as O
from tests import bruteforce as B

GOLD = os.path.join(os.path.dirname(__file__), "golden", "graphs")


def fixture(name, weighted=False, symmetrize=True):
    u, v, w = B.read_el(os.path.join(GOLD, name), weighted)
    return O.make_graph(u, v, w, symmetrize=symmetrize), (u, v, w)


# ------------------------------------------------------------- RNG / generator
def test_mt19937_standard_kat():
    # [rand.predef]: the 10000th output of a default-constructed mt19937.
    assert int(O.mt19937(5489, 10000)[-1]) == 4123659995
    py = B.MT19937(27491095)
    ref = O.mt19937(27491095, 700)
    assert [py() for _ in range(700)] == ref.tolist()


def test_splitmix64_published_vector():
    # SplitMix64 from state 0: first output 0xe220a8397b1dcdaf.
    assert B.splitmix64_stream(0, 1)[0] == 0xE220A8397B1DCDAF


def _uniform_el_python(scale, degree):
    """MakeUniformEL (gapbs/generator.h:101-117) restated from the MT stream."""
    n = 1 << scale
    m = n * degree
    block = 1 << 18
    us, vs = [], []
    for b in range(0, m, block):
        rng = B.MT19937(27491095 + b // block)
        for _ in range(b, min(b + block, m)):
            # Edge(udist(), udist()): the reference's GCC build evaluates the
            # arguments right to left, so the destination takes the first draw
            # (the reference compiled here shows it; tests/golden/ref_vectors.json)
            vs.append(rng() % n)   # n is a power of two: no rejection cutoff
            us.append(rng() % n)
    return np.array(us, np.int32), np.array(vs, np.int32)


def test_uniform_generator_matches_mt_restatement():
    u, v = O.generate_el(6, 4, uniform=True)
    pu, pv = _uniform_el_python(6, 4)
    assert np.array_equal(u, pu) and np.array_equal(v, pv)


def _rmat_unpermuted_python(scale, degree):
    n = 1 << scale
    m = n * degree
    A = int(0.57 * 4294967295)
    Bq = int(0.19 * 4294967295)
    Cq = int(0.19 * 4294967295)
    edges = []
    block = 1 << 18
    for b in range(0, m, block):
        rng = B.MT19937(27491095 + b // block)
        for _ in range(b, min(b + block, m)):
            s = d = 0
            for _ in range(scale):
                r = rng()
                s <<= 1
                d <<= 1
                if r < A + Bq:
                    if r > A:
                        d += 1
                else:
                    s += 1
                    if r > A + Bq + Cq:
                        d += 1
            edges.append((s, d))
    return edges


def test_rmat_generator_is_permuted_python_rmat():
    scale, degree = 7, 4
    u, v = O.generate_el(scale, degree)
    raw = _rmat_unpermuted_python(scale, degree)
    # PermuteIDs applies one vertex permutation to both endpoints: recover it
    # from edge 0.. and check every edge maps consistently.
    perm = {}
    for (s, d), a, b in zip(raw, u.tolist(), v.tolist()):
        for x, y in ((s, a), (d, b)):
            assert perm.setdefault(x, y) == y
    assert len(set(perm.values())) == len(perm)  # injective


def test_rmat_quadrant_probabilities():
    # scale 1: each edge is one quadrant draw; P(a)=.57, P(b)=P(c)=.19, P(d)=.05
    u, v = O.generate_el(1, 8192)
    same = np.sum(u == v)
    diag0 = np.sum((u == v) & (u == u[0] if u[0] == v[0] else u == 1 - u[0]))
    frac_diag = same / len(u)
    assert abs(frac_diag - 0.62) < 0.02
    counts = sorted([np.sum((u == a) & (v == b)) / len(u) for a in (0, 1) for b in (0, 1)])
    assert abs(counts[0] - 0.05) < 0.01 and abs(counts[3] - 0.57) < 0.02
    assert abs(counts[1] - 0.19) < 0.02 and abs(counts[2] - 0.19) < 0.02
    del diag0


def test_insert_weights_matches_mt_restatement():
    m = 1000
    w = O.insert_weights(m)
    rng = B.MT19937(27491095)
    exp = []
    for _ in range(m):
        r = rng()
        while r >= 4294967295:
            r = rng()
        exp.append(r % 255 + 1)
    assert w.tolist() == exp
    assert w.min() >= 1 and w.max() <= 255


def test_generator_block_reseed_beyond_first_block():
    # 2^18 edges per block: scale 15 x 16 = 2 blocks (pins the reseeding).
    u, v = O.generate_el(15, 16, uniform=True)
    rng = B.MT19937(27491095 + 1)
    first_v, first_u = rng() % 32768, rng() % 32768  # destination drawn first
    assert (int(u[1 << 18]), int(v[1 << 18])) == (first_u, first_v)


# ------------------------------------------------------------------ CSR build
@pytest.mark.parametrize("sym", [True, False])
@pytest.mark.parametrize("weighted", [False, True])
def test_build_csr_matches_set_semantics(sym, weighted):
    u, v, w = B.random_graph(60, 400, seed=3, weighted=weighted)
    n = int(max(u.max(), v.max())) + 1
    g = O.make_graph(u, v, w, symmetrize=sym, n=n)
    off, nbr, wgt = B.csr_from_sets(B.adjacency_sets(n, u, v, w, symmetrize=sym))
    assert np.array_equal(g.off, off) and np.array_equal(g.nbr, nbr)
    if weighted:
        assert np.array_equal(g.wgt, wgt)
    if not sym:  # in-CSR = transpose
        ioff, inbr, iwgt = B.csr_from_sets(B.adjacency_sets(n, v, u, w, symmetrize=False))
        assert np.array_equal(g.in_off, ioff) and np.array_equal(g.in_nbr, inbr)


def test_kron_invariants():
    g = O.kron(10)
    assert g.n <= 1024
    for x in range(g.n):
        row = g.nbr[g.off[x]:g.off[x + 1]]
        assert np.all(np.diff(row) > 0) and not np.any(row == x)
    # symmetric
    pairs = set()
    for x in range(g.n):
        for y in g.nbr[g.off[x]:g.off[x + 1]]:
            pairs.add((x, int(y)))
    assert all((b, a) in pairs for a, b in pairs)


# --------------------------------------------------------------- relabel / ids
def _edges(g):
    out = []
    for x in range(g.n):
        for k in range(g.off[x], g.off[x + 1]):
            out.append((x, int(g.nbr[k]), int(g.wgt[k]) if g.wgt is not None else 1))
    return sorted(out)


def test_relabel_reference_fixtures():
    """bench/tests/test_relabel_weights.cc:23-134 fixtures and mapping {2,0,3,1}."""
    mapping = np.array([2, 0, 3, 1], np.int32)
    und = O.Graph(n=4, off=np.array([0, 2, 4, 6, 8], np.int64),
                  nbr=np.array([1, 2, 0, 3, 0, 3, 1, 2], np.int32),
                  wgt=np.array([5, 7, 5, 11, 7, 13, 11, 13], np.int32))
    r = O.relabel(und, mapping)
    exp = sorted((int(mapping[a]), int(mapping[b]), w) for a, b, w in _edges(und))
    assert _edges(r) == exp
    # explicit rows: new row 0 = old row 1 -> {0->2 (5), 3->1 (11)} sorted
    assert r.nbr[r.off[0]:r.off[1]].tolist() == [1, 2]
    assert r.wgt[r.off[0]:r.off[1]].tolist() == [11, 5]
    d = O.Graph(n=4, off=np.array([0, 2, 2, 3, 4], np.int64),
                nbr=np.array([1, 2, 1, 0], np.int32), wgt=np.array([5, 7, 9, 11], np.int32),
                in_off=np.array([0, 1, 3, 4, 4], np.int64),
                in_nbr=np.array([3, 0, 2, 0], np.int32), in_wgt=np.array([11, 5, 9, 7], np.int32),
                directed=True)
    rd = O.relabel(d, mapping)
    assert _edges(rd) == sorted((int(mapping[a]), int(mapping[b]), w) for a, b, w in _edges(d))
    assert r.org_ids.tolist() == [1, 3, 0, 2]


def test_relabel_random_graph_roundtrip():
    g = O.kron(9, weighted=True)
    ids = O.mapping_random(g.n)
    r = O.relabel(g, ids)
    inv = np.argsort(ids).astype(np.int32)
    back = O.relabel(r, inv)
    assert np.array_equal(back.off, g.off) and np.array_equal(back.nbr, g.nbr)
    assert np.array_equal(back.wgt, g.wgt)
    assert np.array_equal(back.org_ids, np.arange(g.n, dtype=np.int32))
    # degree sequence permuted
    assert np.array_equal(np.diff(r.off)[ids], np.diff(g.off))


def test_relabel_rejects_non_permutation():
    g = O.kron(6)
    bad = np.arange(g.n, dtype=np.int32)
    bad[1] = 0
    with pytest.raises(ValueError):
        O.relabel(g, bad)
    assert O.validate_permutation(bad) == -2
    bad[1] = g.n
    assert O.validate_permutation(bad) == -1


def test_random_mapping_reference_kat():
    """scripts/test/test_lightweight_baselines.py:14-21,76: the .lo written after
    -o 1 on the 8-vertex two-star graph is (2,6,0,3,4,1,5,7)."""
    u = np.array([6, 6, 6, 7, 7, 7], np.int32)
    v = np.array([0, 1, 2, 3, 4, 5], np.int32)
    g = O.make_graph(u, v)
    r = O.relabel(g, O.mapping_random(g.n))
    assert tuple(r.org_ids.tolist()) == (2, 6, 0, 3, 4, 1, 5, 7)


def test_random_mapping_matches_splitmix_fisher_yates():
    n = 50
    stream = iter(B.splitmix64_stream(0, 1000))
    perm = list(range(n))
    for rem in range(n, 1, -1):
        thr = ((1 << 64) - rem) % rem
        x = next(stream)
        while x < thr:
            x = next(stream)
        j = x % rem
        perm[rem - 1], perm[j] = perm[j], perm[rem - 1]
    assert O.mapping_random(n).tolist() == perm


def test_sort_mapping_is_stable_descending():
    g = O.kron(8)
    ids = O.mapping_sort(g)
    deg = np.diff(g.off)
    order = np.empty(g.n, np.int64)
    order[ids] = np.arange(g.n)
    assert np.all(np.diff(deg[order]) <= 0)
    for a, b in zip(order[:-1], order[1:]):
        if deg[a] == deg[b]:
            assert a < b


def test_mapping_fingerprint_fnv():
    # builder.h:1400 seeds FNV-1a with 1469598103934665603, the published
    # 64-bit offset basis 14695981039346656037 with its last digit dropped;
    # the prime is the published 0x100000001b3.
    basis, prime = 1469598103934665603, 0x100000001B3
    assert O.mapping_fingerprint(np.zeros(0, np.int32)) == f"{basis:016x}"
    ids = np.array([3, 1, 2, 0], np.int32)
    h = basis
    for i, x in enumerate(ids.tolist()):
        h = ((h ^ i) * prime) & ((1 << 64) - 1)
        h = ((h ^ x) * prime) & ((1 << 64) - 1)
    assert O.mapping_fingerprint(ids) == f"{h:016x}"


def test_source_picker_is_ordering_invariant():
    """bench/tests/test_relabel_weights.cc:167-206."""
    g = O.kron(10)
    o1, i1 = O.source_picker(g, 16)
    ids = O.mapping_random(g.n)
    r = O.relabel(g, ids)
    o2, i2 = O.source_picker(r, 16)
    assert np.array_equal(o1, o2) and np.array_equal(i1, o1)
    assert np.array_equal(i2, ids[o1])
    assert np.all(np.diff(g.off)[i1] > 0)
    o3, _ = O.source_picker(g, 16, repeats=4)
    assert np.array_equal(o3, np.repeat(o1[:4], 4))


# -------------------------------------------------------------------- kernels
FIXTURES = ["tiny.el", "path.el", "star.el", "disconnected.el", "isolated.el", "dangling.el"]


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("sym", [True, False])
def test_bfs_on_fixtures(name, sym):
    g, _ = fixture(name, symmetrize=sym)
    for s in range(g.n):
        if g.off[s + 1] == g.off[s]:
            continue
        depth = O.bfs_depths(g, s)
        assert np.array_equal(depth, B.bfs_depths(g.n, g.off, g.nbr, s))
        for alpha, beta in ((15, 18), (1, 1), (10 ** 6, 18)):
            p = O.bfs_do(g, s, alpha, beta)
            assert O.bfs_verify(g, s, p)


def test_bfs_path_depths_closed_form():
    g, _ = fixture("path.el")
    assert O.bfs_depths(g, 0).tolist() == [0, 1, 2, 3, 4, 5]


def test_bfs_verifier_rejects_wrong_parent():
    g = O.kron(8)
    s = int(O.source_picker(g, 1)[1][0])
    p = O.bfs_do(g, s)
    bad = p.copy()
    x = int(np.argmax(O.bfs_depths(g, s)))
    bad[x] = x
    assert not O.bfs_verify(g, s, bad)


@pytest.mark.parametrize("threads", [1, 4])
def test_bfs_kron_parents_verify(threads):
    g = O.kron(12)
    _, srcs = O.source_picker(g, 4)
    for s in srcs:
        p = O.bfs_do(g, int(s), threads=threads)
        assert O.bfs_verify(g, int(s), p)
        d = O.bfs_depths(g, int(s))
        assert np.array_equal(p >= 0, d >= 0)


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("sym", [True, False])
def test_cc_on_fixtures(name, sym):
    g, (u, v, _) = fixture(name, symmetrize=sym)
    # weak components = components of the symmetrized graph
    gs = O.make_graph(u, v, symmetrize=True, n=g.n)
    exp = B.components_min_label(g.n, gs.off, gs.nbr)
    assert np.array_equal(O.cc_afforest(g), exp)
    if sym:
        assert np.array_equal(O.cc_sv(g), exp)
    assert O.cc_verify(g, exp)


def test_cc_disconnected_count():
    g, _ = fixture("disconnected.el")
    assert len(np.unique(O.cc_afforest(g))) == 3


@pytest.mark.parametrize("threads", [1, 4])
def test_cc_kron_min_labels(threads):
    g = O.kron(12)
    exp = B.components_min_label(g.n, g.off, g.nbr)
    assert np.array_equal(O.cc_afforest(g, threads=threads), exp)
    assert np.array_equal(O.cc_sv(g, threads=threads), exp)
    bad = exp.copy()
    bad[np.argmax(np.diff(g.off))] = -5
    assert not O.cc_verify(g, bad)


@pytest.mark.parametrize("delta", [1, 2, 8, 32, 128])
def test_sssp_matches_floyd_warshall(delta):
    u, v, w = B.random_graph(40, 160, seed=11, weighted=True)
    g = O.make_graph(u, v, w, symmetrize=False, n=40)
    D, INF = B.floyd_warshall(g.n, g.off, g.nbr, g.wgt)
    for s in range(0, 40, 7):
        if g.off[s + 1] == g.off[s]:
            continue
        exp = np.where(D[s] >= INF, O.K_DIST_INF, D[s]).astype(np.int32)
        assert np.array_equal(O.dijkstra(g, s), exp)
        assert np.array_equal(O.sssp_delta(g, s, delta, threads=4), exp)


def test_sssp_weighted_path_fixture():
    g, _ = fixture("weighted_path.wel", weighted=True, symmetrize=False)
    assert O.dijkstra(g, 0).tolist() == [0, 1, 3, 6, 10, 15]
    assert O.sssp_delta(g, 0, 2).tolist() == [0, 1, 3, 6, 10, 15]


def test_sssp_fingerprints_ordering_invariant():
    g = O.kron(10, weighted=True)
    s_orig = int(O.source_picker(g, 1)[0][0])
    d = O.sssp_delta(g, s_orig, 8, threads=4)
    ids = O.mapping_random(g.n)
    r = O.relabel(g, ids)
    d2 = O.sssp_delta(r, int(ids[s_orig]), 8, threads=4)
    assert O.weight_checksum(g) == O.weight_checksum(r)
    assert O.distance_fingerprint(g, d) == O.distance_fingerprint(r, d2)
    assert np.array_equal(d2[ids], d)


@pytest.mark.parametrize("name", ["tiny.el", "path.el", "star.el", "disconnected.el"])
def test_bc_on_fixtures(name):
    g, _ = fixture(name)
    srcs = [x for x in range(g.n) if g.off[x + 1] > g.off[x]][:4]
    exp = B.brandes(g.n, g.off, g.nbr, srcs)
    for got in (O.bc(g, srcs), O.bc_serial(g, srcs), O.bc_fp64(g, srcs, threads=2)):
        assert np.allclose(got, exp, rtol=1e-5, atol=1e-6)


def test_bc_path_closed_form():
    g, _ = fixture("path.el")
    # from source 0 on a 6-path, delta(k) = number of vertices beyond k.
    assert np.allclose(O.bc_serial(g, [0]), np.array([5, 4, 3, 2, 1, 0]) / 5.0)
    assert np.array_equal(O.bc_fp64(g, [0]), np.array([5, 4, 3, 2, 1, 0]) / 5.0)


def test_bc_kron_parallel_matches_serial():
    g = O.kron(11)
    _, srcs = O.source_picker(g, 4)
    a = O.bc(g, srcs, threads=4)
    b = O.bc_serial(g, srcs)
    assert all(O.bc_within_tolerance(float(x), float(y)) for x, y in zip(b, a))
    exp = B.brandes(g.n, g.off, g.nbr, srcs.tolist())
    assert np.allclose(b, exp, rtol=1e-4, atol=1e-6)
    assert np.allclose(O.bc_fp64(g, srcs, threads=4), exp, rtol=1e-12, atol=1e-14)


def test_bc_tolerance_reference_asserts():
    # bc.cc:61-62 static_asserts
    assert O.bc_within_tolerance(1.0, 1.0 - 2.4e-7)
    assert not O.bc_within_tolerance(0.5, 0.501)


@pytest.mark.parametrize("name", ["tiny.el", "star.el", "dangling.el"])
def test_pr_fp64_matches_dense_solve(name):
    g, _ = fixture(name)
    exact = B.pagerank_dense(g.n, g.off, g.nbr)
    assert np.allclose(O.pr_fp64(g), exact, rtol=1e-10)


def test_pr_float_converges_to_fp64():
    g = O.kron(11)
    ref = O.pr_fp64(g)
    for fn in (O.pr_gs, O.pr_spmv):
        s, iters, err = fn(g, max_iters=100, tol=1e-9, fixed=False)
        assert np.abs(s - ref).sum() / ref.sum() < 1e-4
    s, iters, err = O.pr_gs(g)
    assert iters <= 20 and O.pr_residual(g, s) < 1e-4
    s, iters, err = O.pr_spmv(g, max_iters=5, fixed=True)
    assert iters == 5


def test_pr_dense_random_digraph():
    u, v, _ = B.random_graph(30, 90, seed=5)
    g = O.make_graph(u, v, symmetrize=False, n=30)
    exact = B.pagerank_dense(g.n, g.off, g.nbr)
    assert np.allclose(O.pr_fp64(g), exact, rtol=1e-9)
    s, _, _ = O.pr_spmv(g, max_iters=200, tol=1e-12)
    assert np.allclose(s, exact, rtol=1e-4)


# -------------------------------------------------------------------- Tier-0
def test_tier0_features_python_restatement():
    g = O.kron(12)
    f = O.tier0_features(g)
    n = g.n
    k = max(1024, min(int(np.sqrt(n)), 8192))
    k = min(k, n)
    nodes = [((2 * i + 1) * n) // (2 * k) % n for i in range(k)]
    deg = np.diff(g.off)[nodes].astype(np.float64)
    assert f["avg_degree"] == pytest.approx(deg.mean())
    assert f["degree_cv"] == pytest.approx(deg.std(ddof=1) / deg.mean())
    top = sorted(zip(-deg, nodes))[: max(1, k // 10)]
    assert f["hub_concentration"] == pytest.approx(-sum(t[0] for t in top) / deg.sum())
    span = []
    win = 0
    for x in nodes:
        for y in g.nbr[g.off[x]:g.off[x + 1]]:
            span.append(abs(int(y) - x) / (n - 1))
            win += abs(int(y) // 8 - x // 8) <= 1
    assert f["normalized_edge_span"] == pytest.approx(np.mean(span))
    assert f["window_neighbor_overlap"] == pytest.approx(win / len(span))


def test_frozen_rule_truth_table():
    # reorder_adaptive.h:302-333
    assert O.frozen_rule(1, 1, 5000, 30, 1.0, 1.0)
    assert not O.frozen_rule(1, 3, 5000, 30, 1.0, 1.0)         # reuse > 2
    assert not O.frozen_rule(1, 1, 999, 30, 1.0, 1.0)          # too small
    assert not O.frozen_rule(1, 1, 5000, 30, 1.0, 3.3)         # wsr > 3.2
    assert O.frozen_rule(1, 1, 5000, 100, 1.0, 0.5)            # dense but wsr small
    assert not O.frozen_rule(1, 1, 5000, 100, 1.0, 0.9)
    assert O.frozen_rule(2, 2, 5000, 100, 9.0, 3.0)            # cv > 8
    assert not O.frozen_rule(6, 1, 5000, 30, 1.0, 1.0)         # TC unsupported
    assert O.modeled_property_bytes(1, 10, 100) == 80
    assert O.modeled_property_bytes(4, 10, 100) == 440
    assert O.modeled_property_bytes(5, 10, 100) == 10 * 20 + 13


# ---- selector MLP oracle pinned by finite differences --------------------
def test_mlp_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    d, B = 5, 9
    P = 64 * d + 64 + 64 * 64 + 64 + 64 + 1
    p = rng.normal(0, 0.3, P)
    X = rng.normal(0, 1, (B, d))
    y = rng.normal(0, 1, B)
    loss, g = O.mlp_grad(p, d, X, y)
    assert np.isclose(loss, ((O.mlp_forward(p, d, X) - y) ** 2).mean())
    idx = rng.choice(P, 60, replace=False)
    idx = np.concatenate([idx, [P - 1, P - 2, 0, 64 * d]])  # b3, w3, W1, b1 always
    for i in idx:
        e = np.zeros(P)
        e[i] = 1e-6
        num = (O.mlp_grad(p + e, d, X, y)[0] - O.mlp_grad(p - e, d, X, y)[0]) / 2e-6
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), i


def test_mlp_oracle_known_network():
    # a network whose hidden units copy a positive input: y = sum(x) + b3
    d = 3
    W1 = np.zeros((64, d)); W1[:d] = np.eye(d)
    W2 = np.zeros((64, 64)); W2[:d, :d] = np.eye(d)
    w3 = np.zeros(64); w3[:d] = 1
    p = np.concatenate([W1.ravel(), np.zeros(64), W2.ravel(), np.zeros(64), w3, [0.5]])
    X = np.array([[1.0, 2.0, 3.0], [0.5, 0.0, 4.0]])
    assert np.allclose(O.mlp_forward(p, d, X), [6.5, 5.0])
