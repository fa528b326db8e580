"""find_models in the CPU oracle (matcher.cpp:1077-1124, multi.hpp:14-119).

The reference holds no find_models tests or fixtures, so the oracle's
find_models is pinned by closed forms of the multi-model overlap and by the
selection rules on scenes whose single-model answers are known: every
instance is strong, so the query cache's refinement rules never come into
play, and find_models must reduce to find() per model, to the best match
overall under a global quota, to one match per location under the
cross-model overlap, and to duplicated matches for a repeated handle.
"""
import math

import numpy as np
import pytest

E = math.radians(16)


@pytest.fixture(scope="module")
def multi(oracle):
    from mvkit_amd import synth
    s = synth.small_multi()
    models = [oracle.Model(t, levels=3, start=-E, extent=2 * E) for t in s.templates]
    return s, models


def q(**kw):
    base = dict(start=-E, extent=2 * E, min_score=0.7, max_overlap=0.5, num_matches=0)
    base.update(kw)
    return base


def test_box_overlap_closed_forms(oracle):
    """Intersection over the smaller rectangle (multi.hpp:64-81)."""
    g = (40, 20, 9.5, 19.5)
    assert oracle.box_overlap((50, 50, 0), g, (50, 50, 0), g) == pytest.approx(1)
    assert oracle.box_overlap((50, 50, 0), g, (50, 70, 0), g) == pytest.approx(0.5)
    assert oracle.box_overlap((50, 50, 0), g, (55, 60, 0), g) == pytest.approx(30 * 15 / 800)
    assert oracle.box_overlap((50, 50, 0), g, (50, 150, 0), g) == 0
    # a small rectangle inside a large one covers all of the smaller area
    small = (10, 10, 4.5, 4.5)
    big = (60, 60, 29.5, 29.5)
    assert oracle.box_overlap((50, 50, 0.3), small, (50, 50, 0), big) == pytest.approx(1)
    assert oracle.box_overlap((50, 50, 0), big, (50, 50, 0.3), small) == pytest.approx(1)
    # a square turned by a quarter turn about its centre is itself
    sq = (30, 30, 14.5, 14.5)
    assert oracle.box_overlap((50, 50, 0), sq, (50, 50, math.pi / 2), sq) == pytest.approx(1)
    # rotated rectangles of two sizes: Monte Carlo of the rasterised intersection
    rng = np.random.default_rng(11)
    ga, gb = (40, 20, 9.5, 19.5), (24, 30, 14.5, 11.5)
    a, b = (50.0, 50.0, 0.3), (56.0, 58.0, -0.5)
    pts = rng.uniform([10, 10], [90, 100], size=(400000, 2))

    def inside(p, pose, geo):
        w, h, orow, ocol = geo
        r, c, t = pose
        dy, dx = p[:, 0] - r, p[:, 1] - c
        u = math.cos(t) * dx - math.sin(t) * dy
        v = math.sin(t) * dx + math.cos(t) * dy
        return (u >= -ocol - .5) & (u <= w - ocol - .5) & (v >= -orow - .5) & (v <= h - orow - .5)

    inter = (inside(pts, a, ga) & inside(pts, b, gb)).mean() * 80 * 90
    frac = inter / min(ga[0] * ga[1], gb[0] * gb[1])
    assert oracle.box_overlap(a, ga, b, gb) == pytest.approx(frac, abs=0.01)


def test_single_model_reduces_to_find(oracle, multi):
    s, models = multi
    for m in models:
        want = m.find(s.image, -E, 2 * E, 0.7, 0, 0.5)
        got = oracle.find_models([m], s.image, [q()], selection_matches=0, selection_overlap=0.5)
        assert len(want) == 2 and got.shape == (2, 5)
        assert np.array_equal(got[:, 4], [0, 0])
        assert np.allclose(got[:, :4], want, atol=1e-9, rtol=0)


def test_per_model_controls_reduce_to_finds(oracle, multi):
    """Per-model NumMatches / MaxOverlap: each model alone, merged in
    (score desc, model) order; every pasted instance found by its own model."""
    s, models = multi
    got = oracle.find_models(models, s.image, [q(num_matches=2), q(num_matches=2)], -1, -1)
    want = []
    for k, m in enumerate(models):
        for r in m.find(s.image, -E, 2 * E, 0.7, 2, 0.5):
            want.append((*r, k))
    want.sort(key=lambda x: (-x[3], x[4]))
    assert got.shape == (4, 5)
    assert np.allclose(got, np.array(want), atol=1e-9, rtol=0)
    for k, r, c, a in s.truth:
        hit = [g for g in got if g[4] == k and math.hypot(g[0] - r, g[1] - c) < 1.0]
        assert len(hit) == 1 and abs(hit[0][2] - a) < 0.02


def test_global_quota_takes_the_first_candidate(oracle, multi):
    """A one-value NumMatches stops at the first accepted matches in the
    merged order of the prepared (unrefined) scores: here the top prepared
    candidate of all models, whose refined score is not the highest."""
    s, models = multi
    every, prep = oracle.find_models(models, s.image, [q(), q()], 0, 0.5, prepared=True)
    best = oracle.find_models(models, s.image, [q(num_matches=1), q(num_matches=1)], 1, 0.5)
    assert every.shape[0] == 4 and best.shape == (1, 5)
    assert any(np.array_equal(best[0], e) for e in every)
    tops = [(p[0, 3], -k, k, p[0]) for k, p in enumerate(prep)]
    _, _, k_top, p_top = max(tops, key=lambda t: (t[0], t[1]))
    assert best[0, 4] == k_top
    assert math.hypot(best[0, 0] - p_top[0], best[0, 1] - p_top[1]) < 1.0


def test_cross_model_overlap_and_model_order(oracle, multi):
    """Two model objects of one template: a one-value MaxOverlap suppresses
    the second model's identical matches (ties go to the lower model index);
    per-model MaxOverlap keeps both."""
    s, models = multi
    twin = oracle.Model(s.templates[0], levels=3, start=-E, extent=2 * E)
    pair = [models[0], twin]
    shared = oracle.find_models(pair, s.image, [q(), q()], 0, 0.5)
    alone = oracle.find_models([models[0]], s.image, [q()], 0, 0.5)
    assert np.array_equal(shared[:, 4], np.zeros(len(shared)))
    assert np.allclose(shared[:, :4], alone[:, :4], atol=1e-9, rtol=0)
    both = oracle.find_models(pair, s.image, [q(), q()], -1, -1)
    assert len(both) == 2 * len(alone)
    assert np.array_equal(both[:, 4], np.tile([0, 1], len(alone)))
    assert np.allclose(both[0::2, :4], both[1::2, :4], atol=0, rtol=0)


def test_two_stage_quota(oracle, multi):
    """Per-model quotas with a one-value MaxOverlap: quotas first, then the
    cross-model overlap, which does not refill a quota (matcher.cpp:1108-1112)."""
    s, models = multi
    twin = oracle.Model(s.templates[0], levels=3, start=-E, extent=2 * E)
    pair = [models[0], twin]
    got = oracle.find_models(pair, s.image, [q(num_matches=1), q(num_matches=1)], -1, 0.5)
    assert got.shape == (1, 5) and got[0, 4] == 0
    # without the two stages the twin's quota would refill with its second instance
    one_stage = oracle.find_models(pair, s.image, [q(num_matches=1), q(num_matches=1)], -1, 0.5,
                                   per_model_quota=False)
    assert one_stage.shape == (2, 5) and list(one_stage[:, 4]) == [0, 1]
    assert math.hypot(*(one_stage[0, :2] - one_stage[1, :2])) > 20


def test_repeated_handle_shares_candidates(oracle, multi):
    s, models = multi
    got = oracle.find_models([models[1], models[1]], s.image, [q(), q()], -1, -1)
    assert len(got) == 4
    assert np.array_equal(got[:, 4], [0, 1, 0, 1])
    assert np.array_equal(got[0::2, :4], got[1::2, :4])
