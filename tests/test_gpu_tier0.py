"""Device Tier-0 features (gbk_tier0_features) against the host restatement
and the reference's own values (GPU).

Bars: every feature bit-exact to the host (which equals the reference's
ComputeTier0SampledGraphFeatures, tests/test_oracle_ref.py) except the
normalised edge span, which sums exact integer spans and divides once:
1e-12 relative.  Default sampling, explicit sizes and seeds, up to the 8192
maximum, directed and undirected graphs."""
import json
import os

import numpy as np
import pytest

import oracle as O
from graphbrew_amd import host as H

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _directed():
    rng = np.random.default_rng(5)
    n, m = 3000, 30000
    u = rng.integers(0, n, m).astype(np.int32)
    v = (rng.integers(0, n, m) * rng.integers(0, n, m) // n).astype(np.int32)
    return O.make_graph(u, v, symmetrize=False, n=n)


GRAPHS = {
    "kron12": lambda: O.kron(12),
    "uniform12": lambda: O.kron(12, uniform=True),
    "kron14": lambda: O.kron(14),
    "kron16": lambda: O.kron(16),
    "directed3000": _directed,
    "star": lambda: O.make_graph(np.zeros(5000, np.int32), np.arange(1, 5001, dtype=np.int32)),
}


def _same(dev, host, what):
    for k in host:
        if k == "normalized_edge_span":
            assert dev[k] == pytest.approx(host[k], rel=1e-12, abs=0), (what, k)
        else:
            assert dev[k] == host[k], (what, k, dev[k], host[k])


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_device_tier0_equals_host(ctx, name):
    g = GRAPHS[name]()
    dg = ctx.upload_graph(g)
    for kw in ({}, {"sample_size": 2000, "seed": 7}, {"sample_size": 8192, "seed": 3},
               {"sample_size": 37, "seed": 1 << 40}):
        if kw.get("sample_size", 0) > g.n:
            continue
        _same(ctx.tier0_features(dg, **kw), H.tier0_features(g.off, g.nbr, **kw), (name, kw))
    dg.free()


def test_device_tier0_equals_reference_vectors(ctx):
    with open(os.path.join(GOLD, "ref_vectors.json")) as f:
        V = json.load(f)["tier0"]
    for name, rec in V.items():
        g = GRAPHS[name]()
        dg = ctx.upload_graph(g)
        for key, kw in (("default", {}), ("s2000_seed7", {"sample_size": 2000, "seed": 7})):
            dev = ctx.tier0_features(dg, **kw)
            _same(dev, dict(zip(H.TIER0_NAMES, rec[key])), (name, key))
        dg.free()


def test_device_tier0_rejects_oversized_sample(ctx):
    from graphbrew_amd import gbk
    dg = ctx.upload_graph(O.kron(14))
    with pytest.raises(gbk.GbkError):
        ctx.tier0_features(dg, sample_size=8193)
    dg.free()
