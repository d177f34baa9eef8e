"""The adaptive max-iter rule (math_core/engine.py:1051-1060) on the CPU, and
the 128x72 probe pass on the GPU against the same rule over the oracle's map
of the probe view."""
from decimal import Decimal

import numpy as np
import pytest

from fractura_amd import engine


def test_rule_keeps_the_base_when_nothing_escaped():
    assert engine.get_adaptive_max_iter_from_probe(np.full((72, 128), -1.0, np.float32), 777) == 777


def test_rule_closed_forms():
    rule = engine.get_adaptive_max_iter_from_probe
    # One escaped value v: p98 = max = v -> max(1.35 v, 1.2 v) = 1.35 v.
    a = np.full(100, -1.0)
    a[3] = 1000.0
    assert rule(a, 5) == 1350
    # Small values are lifted to 100.
    assert rule(np.array([10.0, 20.0]), 5) == 100
    # A lone outlier: 1.2 x max beats 1.35 x p98.
    b = np.concatenate([np.full(1000, 100.0), [5000.0]])
    p98 = float(np.percentile(b, 98))
    assert rule(b, 5) == int(max(p98 * 1.35, 5000.0 * 1.2)) == 6000
    # Clamped to the global maximum.
    assert rule(np.array([1e12]), 5) == engine.GLOBAL_MAX_ITER


@pytest.mark.gpu
def test_probe_pass_matches_the_rule_on_the_oracle_map():
    from oracle import oracle

    cx, cy, w = Decimal("-0.7435"), Decimal("0.1314"), Decimal("0.002")
    got = engine.probe_adaptive_max_iter(cx, cy, w, 500)
    plan = engine.plan_view(128, 72, max(1000, 2 * 500), cx, cy, w)
    assert plan.engine == "fp64"
    m = oracle.render("fp64", plan.view)
    want = engine.get_adaptive_max_iter_from_probe(m.iter_map, 500)
    assert got == want
    assert got != 500  # the view escapes, so the rule moved the value
