"""Learned ordering selector (GPU): trained by the library's own kernels on
the committed timings (graphbrew_amd/selector_corpus.json, measured by
graphbrew_amd.selector.collect on this build, mapping costs as the drivers
pay them), judged on held-out graphs by measured end-to-end regret against
the best candidate, the original order and the frozen rule's deployable pick
(DBG: its fallback branch in builds without Rabbit order, and its stand-in
where it asks for the Leiden-Gorder8 composition, which is not produced
here; reorder_adaptive.h:302-383)."""
import json
import os

import numpy as np
import pytest

from graphbrew_amd import selector as S

pytestmark = pytest.mark.gpu
CORPUS = os.path.join(os.path.dirname(S.MODEL_PATH), "selector_corpus.json")


@pytest.fixture(scope="module")
def corpus():
    with open(CORPUS) as f:
        return json.load(f)


def test_trained_network_beats_the_frozen_rule(ctx, corpus):
    train, test = S.split(corpus)
    sel = S.Selector()
    curve = sel.fit(ctx, train)
    assert curve[-1] < 0.2 * curve[0]
    ev = S.evaluate(ctx, sel, test)
    m, fb, orig = ev["model"], ev["fallback"], ev["original"]
    # what the network gains over the rule it replaces and over the original
    # order, at each reuse count (retraining sums gradients with atomics, so
    # a retrained network may differ from the committed one in the last
    # digits: 0.1 point of slack here, none for the committed model below)
    for r in ("1", "2"):
        br = ev["by_reuse"][r]
        assert br["model"]["mean_regret"] < 0.1 * br["fallback"]["mean_regret"]
        assert br["model"]["mean_regret"] <= br["original"]["mean_regret"] + 0.001, r
    assert m["within_10pct"] >= 0.95


def test_committed_model_reproduces_its_evaluation(ctx, corpus):
    sel = S.Selector.load()
    assert sel.params is not None and sel.params.shape[1] == 64 * S.IN_DIM + 64 + 64 * 64 + 64 + 64 + 1
    _, test = S.split(corpus)
    ev = S.evaluate(ctx, sel, test)
    rec = sel.meta["held_out"]
    assert ev["picks"] == rec["picks"]
    assert abs(ev["model"]["mean_regret"] - rec["model"]["mean_regret"]) < 1e-9
    # on the held-out graphs the committed model is no worse than the
    # original order and far better than the frozen rule at reuse 1 and 2
    for r in ("1", "2"):
        br = ev["by_reuse"][r]
        assert br["model"]["mean_regret"] <= br["original"]["mean_regret"], r
        assert br["model"]["mean_regret"] < br["fallback"]["mean_regret"], r
        assert br["model"]["within_10pct"] >= 0.95, r


def test_untrained_selector_falls_back_to_dbg(ctx, corpus):
    sel = S.Selector()
    g = corpus[0]
    for k in S.KERNELS:
        assert sel.pick(ctx, g["features"], k, 1) == S.FALLBACK == 5
