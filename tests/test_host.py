"""CPU checks of the product host substrate (libgbkhost.so) against the
oracle: generator streams, readers, mapping producers/files, sources,
Tier-0 features and the frozen rule must be identical."""
import os

import numpy as np
import pytest

import oracle as O
from graphbrew_amd import host as H

GOLD = os.path.join(os.path.dirname(__file__), "golden", "graphs")


@pytest.mark.parametrize("scale,deg,uniform", [(10, 16, False), (12, 8, True), (15, 16, False)])
def test_generator_matches_oracle(scale, deg, uniform):
    u, v = H.generate_el(scale, deg, uniform)
    ou, ov = O.generate_el(scale, deg, uniform)
    assert np.array_equal(u, ou) and np.array_equal(v, ov)


def test_insert_weights_matches_oracle():
    m = (1 << 18) + 12345  # crosses a stream boundary
    assert np.array_equal(H.insert_weights(m), O.insert_weights(m))


@pytest.mark.parametrize("name,weighted", [("tiny.el", False), ("weighted.wel", True),
                                           ("isolated.el", False), ("weighted_isolated.wel", True)])
def test_read_el_fixtures(name, weighted, tmp_path):
    u, v, w = H.read_el(os.path.join(GOLD, name), weighted)
    from tests import bruteforce as B
    bu, bv, bw = B.read_el(os.path.join(GOLD, name), weighted)
    assert np.array_equal(u, bu) and np.array_equal(v, bv)
    if weighted:
        assert np.array_equal(w, bw)
    p = tmp_path / "c.el"
    p.write_text("# comment\n% other\n\n3 4\n5 6 7\n")
    u, v, _ = H.read_el(str(p))
    assert u.tolist() == [3, 5] and v.tolist() == [4, 6]


def test_mappings_match_oracle():
    g = O.kron(12)
    assert np.array_equal(H.mapping_random(g.n), O.mapping_random(g.n))
    assert np.array_equal(H.mapping_sort(g.off), O.mapping_sort(g))
    assert np.array_equal(H.mapping_identity(7), np.arange(7))
    ids = O.mapping_random(g.n)
    assert H.mapping_fingerprint(ids) == O.mapping_fingerprint(ids)
    assert H.validate_permutation(ids) == 0
    bad = ids.copy()
    bad[0] = bad[1]
    assert H.validate_permutation(bad) == -2
    bad[0] = -3
    assert H.validate_permutation(bad) == -1


def test_mapping_files_roundtrip(tmp_path):
    ids = O.mapping_random(1000)
    for ext in (".so", ".lo"):
        p = str(tmp_path / ("map" + ext))
        H.mapping_write(p, ids)
        assert np.array_equal(H.mapping_read(p), ids)
    assert (tmp_path / "map.so").stat().st_size == 4000


def test_source_picker_matches_oracle():
    g = O.kron(12)
    r = O.relabel(g, O.mapping_random(g.n))
    for graph in (g, r):
        for reps in (1, 3):
            ho, hi = H.source_picker(graph.off, 20, graph.org_ids, reps)
            oo, oi = O.source_picker(graph, 20, reps)
            assert np.array_equal(ho, oo) and np.array_equal(hi, oi)


def test_afforest_samples_match_oracle():
    assert np.array_equal(H.afforest_samples(65536), O.afforest_sample_indices(65536))


def test_tier0_and_rule_match_oracle():
    g = O.kron(13)
    hf = H.tier0_features(g.off, g.nbr)
    of = O.tier0_features(g)
    for k in H.TIER0_NAMES:
        assert hf[k] == pytest.approx(of[k], rel=1e-12, abs=1e-15)
    for bench, code in H.BENCH_CODES.items():
        assert H.modeled_property_bytes(bench, 1000, 20000) == O.modeled_property_bytes(code, 1000, 20000)
        for args in ((1, 5000, 30, 1.0, 1.0), (3, 5000, 30, 1.0, 1.0), (1, 5000, 100, 9.0, 3.0)):
            assert H.frozen_rule(bench, *args) == O.frozen_rule(code, *args)


def test_dbg_mapping_matches_oracle_and_closed_form():
    g = O.kron(12)
    assert np.array_equal(H.mapping_dbg(g.off), O.mapping_dbg(g))
    # two stars 6-{0,1,2}, 7-{3,4,5}: avg degree 1, leaves in bucket 1, hubs
    # (degree 3) in bucket 3 -> hubs first, then leaves in id order
    star = O.make_graph(np.array([6, 6, 6, 7, 7, 7], np.int32), np.arange(6, dtype=np.int32))
    assert H.mapping_dbg(star.off).tolist() == [2, 3, 4, 5, 6, 7, 0, 1]


def test_host_verifiers_agree_with_oracle():
    g = O.kron(11)
    gw = O.kron(10, weighted=True)
    _, srcs = O.source_picker(g, 3)
    for s in srcs:
        p = O.bfs_do(g, int(s))
        assert H.verify_bfs(g.off, g.nbr, g.off, g.nbr, int(s), p)
        bad = p.copy()
        far = int(np.argmax(O.bfs_depths(g, int(s))))
        bad[far] = far
        assert not H.verify_bfs(g.off, g.nbr, g.off, g.nbr, int(s), bad)
    c = O.cc_afforest(g)
    assert H.verify_cc(g.off, g.nbr, g.off, g.nbr, False, c)
    c2 = c.copy()
    c2[np.argmax(np.diff(g.off))] = -7
    assert not H.verify_cc(g.off, g.nbr, g.off, g.nbr, False, c2)
    sw = int(O.source_picker(gw, 1)[1][0])
    d = O.dijkstra(gw, sw)
    assert H.verify_sssp(gw.off, gw.nbr, gw.wgt, sw, d) == 0
    d2 = d.copy()
    d2[d2 < O.K_DIST_INF][:1] += 1
    d2[np.flatnonzero(d < O.K_DIST_INF)[1]] += 1
    assert H.verify_sssp(gw.off, gw.nbr, gw.wgt, sw, d2) == 1
    assert H.weight_checksum(gw.off, gw.nbr, gw.wgt) == O.weight_checksum(gw)
    assert H.distance_fingerprint(d) == O.distance_fingerprint(gw, d)
    b = O.bc_serial(g, srcs)
    ok, worst = H.verify_bc(g.off, g.nbr, srcs, b)
    assert ok and worst < 1e-6
    b2 = b.copy()
    b2[np.argmax(b2)] *= 0.9
    assert not H.verify_bc(g.off, g.nbr, srcs, b2)[0]
    s, _, _ = O.pr_gs(g)
    assert H.pr_residual(g.off, g.nbr, s) == pytest.approx(O.pr_residual(g, s), rel=1e-9)


# ---- hub mappings (reorder/reorder_hub.h) ---------------------------------
def _off(degrees):
    return np.concatenate([[0], np.cumsum(degrees)]).astype(np.int64)


def test_hub_mappings_known_answers():
    # degrees 2,100,5,50 (the example of reorder_hub.h:111-115): average 39,
    # hubs 1 and 3; vertex 1 gets id 0, vertex 3 id 1
    off = _off([2, 100, 5, 50])
    assert H.mapping_hubsort(off).tolist() == [3, 0, 2, 1]
    assert H.mapping_hubcluster(off).tolist() == [3, 0, 2, 1]
    # degrees 50,2,100,5: hubs 0 and 2.  HubSort puts 2 first; 2 sits beyond
    # the hub range and its slot 0 belongs to a hub, so 2 is deferred into
    # the free slot of vertex 1.  HubCluster keeps hub order 0,2 and vertex 1
    # takes the slot vertex 2 left.
    off = _off([50, 2, 100, 5])
    assert H.mapping_hubsort(off).tolist() == [1, 2, 0, 3]
    assert H.mapping_hubcluster(off).tolist() == [0, 2, 1, 3]
    assert H.mapping_hubsortdbg(off).tolist() == [1, 2, 0, 3]
    assert H.mapping_hubclusterdbg(off).tolist() == [0, 2, 1, 3]
    # equal degrees: no vertex exceeds the average, every mapping is the identity
    off = _off([3, 3, 3, 3, 3])
    for f in (H.mapping_hubsort, H.mapping_hubcluster, H.mapping_hubsortdbg, H.mapping_hubclusterdbg):
        assert f(off).tolist() == [0, 1, 2, 3, 4]


@pytest.mark.parametrize("scale,uniform", [(10, False), (12, False), (11, True)])
def test_hub_mappings_match_oracle(scale, uniform):
    g = O.kron(scale, uniform=uniform)
    for name in ("hubsort", "hubcluster", "hubsortdbg", "hubclusterdbg"):
        got = getattr(H, "mapping_" + name)(g.off)
        ref = getattr(O, "mapping_" + name)(g.off)
        assert np.array_equal(got, ref), name
        assert H.validate_permutation(got) == 0
        # hubs (degree above m/n) occupy exactly the first ids
        deg = np.diff(g.off)
        hubs = deg > (int(g.off[-1]) // g.n)
        assert set(got[hubs].tolist()) == set(range(int(hubs.sum())))


def test_hubsort_orders_hubs_by_degree():
    g = O.kron(12)
    deg = np.diff(g.off)
    new = H.mapping_hubsortdbg(g.off)
    inv = np.empty_like(new)
    inv[new] = np.arange(g.n)
    h = int((deg > int(g.off[-1]) // g.n).sum())
    d = deg[inv[:h]]
    assert np.all(d[:-1] >= d[1:])
    # the non-hubs keep their relative order
    rest = inv[h:]
    assert np.all(np.diff(rest) > 0)


def test_mapping_dispatch_covers_cheap_orderings():
    g = O.kron(10)
    for a in range(8):
        ids = H.mapping(a, g.off)
        assert H.validate_permutation(ids) == 0
    with pytest.raises(ValueError):
        H.mapping(12, g.off)


# ---- selector corpus and encoding (host side) ------------------------------
def test_selector_corpus_is_complete_and_split_by_family():
    import json
    from graphbrew_amd import selector as S
    with open(os.path.join(os.path.dirname(S.MODEL_PATH), "selector_corpus.json")) as f:
        data = json.load(f)
    assert len(data) >= 96
    for g in data:
        assert len(g["features"]) == 12
        assert sorted(g["orderings"]) == [str(a) for a in range(8)]
        for row in g["orderings"].values():
            assert sorted(row["kernel_ms"]) == sorted(S.KERNELS)
            assert all(v > 0 for v in row["kernel_ms"].values())
        assert g["orderings"]["0"]["map_ms"] == 0 and g["orderings"]["0"]["relabel_ms"] == 0
    train, test = S.split(data)
    fams = lambda part: {g["spec"]["family"] for g in part}  # noqa: E731
    # the four seeded families, and the reference generator's Kronecker graphs
    # at other scales and degrees than the bench graph's
    assert fams(train) == fams(test) == {"rmat", "sbm", "lattice", "powerlaw", "kron"}
    assert not any(g["spec"]["family"] == "kron" and (g["spec"]["scale"], g["spec"]["ef"]) == (22, 16)
                   for g in data)
    assert len(test) == 4 * (len(data) // 16)
    assert len(train) + len(test) == len(data)


def test_selector_encoding_and_samples():
    from graphbrew_amd import selector as S
    feats = np.arange(12, dtype=np.float64)
    x = S.encode(feats, 3, "bfs", 2, np.zeros(12), np.ones(12))
    assert x.shape == (S.IN_DIM,) and S.IN_DIM == 28
    assert np.array_equal(x[:12], feats)
    assert x[12 + 3] == 1 and x[12:20].sum() == 1
    assert x[20 + S.KERNELS.index("bfs")] == 1 and x[20:27].sum() == 1
    assert x[27] == 0.5
    row = {"map_ms": 1.0, "relabel_ms": 2.0, "relabel_w_ms": 3.0, "kernel_ms": {k: 4.0 for k in S.KERNELS}}
    assert S.e2e_ms(row, "pr", 2) == 11.0 and S.e2e_ms(row, "sssp", 1) == 8.0


def test_selector_generators_are_seeded():
    from graphbrew_amd import selector as S
    for spec in S.corpus(8, min_scale=9, max_scale=10):
        (u1, v1), n1 = S.make_edges(spec)
        (u2, v2), n2 = S.make_edges(spec)
        assert n1 == n2 and np.array_equal(u1, u2) and np.array_equal(v1, v2)
        assert u1.min() >= 0 and max(u1.max(), v1.max()) < n1 and not np.any(u1 == v1)


# ------------------------------------------------------- serialized graphs
# Layout (gapbs/writer.h:336-377): bool directed, int64 m, int64 n, out
# offsets[n+1] (int64), m neighbours (int32, or (v, w) int32 pairs in .wsg),
# for a directed graph the in-CSR, then int32 org_ids[n].
import struct

from tests import bruteforce as B


def _sg_bytes(directed, off, nbr, wgt=None, in_off=None, in_nbr=None, in_wgt=None, org=None):
    n, m = len(off) - 1, len(nbr)

    def csr(o, nb, w):
        nb = np.asarray(nb, np.int32)
        body = nb.tobytes() if w is None else np.stack([nb, np.asarray(w, np.int32)], 1).tobytes()
        return np.asarray(o, np.int64).tobytes() + body

    out = struct.pack("<?qq", directed, m, n) + csr(off, nbr, wgt)
    if directed:
        out += csr(in_off, in_nbr, in_wgt)
    return out + np.asarray(org if org is not None else np.arange(n), np.int32).tobytes()


def test_sg_write_matches_layout_and_reads_back(tmp_path):
    g = O.kron(8, weighted=True)
    org = np.random.default_rng(3).permutation(g.n).astype(np.int32)
    p = str(tmp_path / "g.sg")
    H.sg_write(p, g.off, g.nbr, org_ids=org)
    assert open(p, "rb").read() == _sg_bytes(False, g.off, g.nbr, org=org)
    r = H.sg_read(p)
    assert not r["directed"] and np.array_equal(r["off"], g.off) and np.array_equal(r["nbr"], g.nbr)
    assert np.array_equal(r["org_ids"], org)
    pw = str(tmp_path / "g.wsg")
    H.sg_write(pw, g.off, g.nbr, wgt=g.wgt, org_ids=org)
    assert open(pw, "rb").read() == _sg_bytes(False, g.off, g.nbr, g.wgt, org=org)
    rw = H.sg_read(pw, weights=True)
    assert np.array_equal(rw["wgt"], g.wgt) and np.array_equal(rw["nbr"], g.nbr)
    # a .wsg read unweighted drops the weights (reader.h:524-534)
    assert np.array_equal(H.sg_read(pw)["nbr"], g.nbr)
    assert H.sg_info(pw) == {"directed": False, "n": g.n, "m": len(g.nbr), "weighted": True}


def test_sg_directed_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    u = rng.integers(0, 300, 4000).astype(np.int32)
    v = rng.integers(0, 300, 4000).astype(np.int32)
    w = rng.integers(1, 256, 4000).astype(np.int32)
    g = O.make_graph(u, v, w, symmetrize=False, n=300)
    p = str(tmp_path / "d.wsg")
    H.sg_write(p, g.off, g.nbr, g.wgt, g.in_off, g.in_nbr, g.in_wgt)
    assert open(p, "rb").read() == _sg_bytes(True, g.off, g.nbr, g.wgt, g.in_off, g.in_nbr, g.in_wgt)
    r = H.sg_read(p, weights=True)
    for k in ("off", "nbr", "wgt", "in_off", "in_nbr", "in_wgt"):
        assert np.array_equal(r[k], getattr(g, k)), k


def test_sg_generated_weights(tmp_path):
    """A .sg read into a weighted graph: 'unit' gives 1; 'hash' gives
    Mix64((src << 32) ^ dst ^ 0x27491095) % 255 + 1 over original ids, the
    pair ordered for an undirected graph (reader.h:48-59).  Mix64 is
    SplitMix64's output for that state, pinned by its published vector."""
    g = O.kron(7)
    org = np.random.default_rng(5).permutation(g.n).astype(np.int32)
    p = str(tmp_path / "h.sg")
    H.sg_write(p, g.off, g.nbr, org_ids=org)
    assert np.all(H.sg_read(p, weights=True, scheme="unit")["wgt"] == 1)
    got = H.sg_read(p, weights=True, scheme="hash")["wgt"]
    rows = np.repeat(np.arange(g.n), np.diff(g.off))
    for e in range(0, len(g.nbr), 7):
        a, b = int(org[rows[e]]), int(org[g.nbr[e]])
        a, b = min(a, b), max(a, b)
        key = ((a << 32) ^ b ^ 0x27491095) & (2**64 - 1)
        assert got[e] == B.splitmix64_stream(key, 1)[0] % 255 + 1
    # undirected: both directions of an edge carry the same weight
    d = {(int(rows[e]), int(g.nbr[e])): int(got[e]) for e in range(len(g.nbr))}
    assert all(d[(y, x)] == wt for (x, y), wt in d.items())


def test_sg_rejects_bad_files(tmp_path):
    g = O.kron(6)
    p = str(tmp_path / "t.sg")
    H.sg_write(p, g.off, g.nbr)
    data = open(p, "rb").read()
    open(p, "wb").write(data[:-3])                 # truncated org_ids
    with pytest.raises(Exception):
        H.sg_read(p)
    with pytest.raises(Exception):                 # weights into a .sg name
        H.sg_write(str(tmp_path / "x.sg"), g.off, g.nbr, wgt=np.ones(len(g.nbr), np.int32))
    with pytest.raises(Exception):
        H.sg_info(str(tmp_path / "x.el"))
