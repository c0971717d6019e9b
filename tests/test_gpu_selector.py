"""Learned ordering selector (GPU): trained by the library's own kernels on
the committed timings (graphbrew_amd/selector_corpus.json, measured by
graphbrew_amd.selector.collect on this build), judged on held-out graphs by
measured end-to-end regret against the best candidate, the original order
and the frozen rule's deployable pick (DBG; reorder_adaptive.h:370-383)."""
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
    # what the network gains over the rule it replaces, at each reuse count
    for r in ("1", "2"):
        assert ev["by_reuse"][r]["model"]["mean_regret"] < 0.1 * ev["by_reuse"][r]["fallback"]["mean_regret"]
    assert m["within_10pct"] >= 0.9
    # on this corpus the original order is already near-optimal at reuse <= 2;
    # the network must not lose more than one point of regret against it
    assert m["mean_regret"] <= orig["mean_regret"] + 0.01


def test_committed_model_reproduces_its_evaluation(ctx, corpus):
    sel = S.Selector.load()
    assert sel.params is not None and sel.params.shape[1] == 64 * S.IN_DIM + 64 + 64 * 64 + 64 + 64 + 1
    _, test = S.split(corpus)
    ev = S.evaluate(ctx, sel, test)
    rec = sel.meta["held_out"]
    assert ev["picks"] == rec["picks"]
    assert abs(ev["model"]["mean_regret"] - rec["model"]["mean_regret"]) < 1e-9


def test_untrained_selector_falls_back_to_dbg(ctx, corpus):
    sel = S.Selector()
    g = corpus[0]
    for k in S.KERNELS:
        assert sel.pick(ctx, g["features"], k, 1) == S.FALLBACK == 5
