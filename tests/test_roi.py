"""Reduced-domain (ROI) templates and searches on the host: the region and
domain C ABI (region.cpp:64-117, 937-945; image.cpp:267-320), create over a
template's domain (model.cpp:120-171) and the oracle's ROI find
(matcher.cpp:376-393, 505-534, 631-645, 1060).

The reference holds no ROI tests or fixtures.  The pins:
 * canonical runs and domain algebra against numpy masks;
 * create over a rectangular domain equals create from the cropped template
   (same rasters, origin shifted by the crop); over an elliptical domain the
   library's rasters equal the oracle's, and the oracle's model equals the one
   rebuilt from a level 0 built here by hand (centred gray over the mask,
   centroid in long double, model.cpp:138-166);
 * find with a domain: the full rectangle gives the full-image result
   exactly; a rectangle or a small disk around an instance gives that
   instance with the full-image pose and score; an empty domain gives nothing;
   under random sparse domains every match lies in the domain and is a
   full-image match or scores at most as one;
 * find_models with the full rectangle as domain equals find_models without.
"""
import math

import numpy as np
import pytest

from mvkit_amd import synth

E = math.radians(20)


def same_oracle_rasters(a, b):
    assert (a.n_levels, a.n_angles, a.step) == (b.n_levels, b.n_angles, b.step)
    for lv in range(-1, a.n_levels):
        for k in ([0] if lv < 0 else range(a.n_angles)):
            A, B = a.patch(lv, k), b.patch(lv, k)
            for key in ("n", "angle", "orow", "ocol", "rmin", "rmax", "cmin", "cmax", "wsum", "n_runs", "n_wruns"):
                assert A[key] == B[key], (lv, k, key)
            for key in ("rows", "cols", "w"):
                assert np.array_equal(A[key], B[key])


def mask_of(runs, h, w):
    m = np.zeros((h, w), bool)
    for r, a, b in np.asarray(runs).reshape(-1, 3):
        if 0 <= r < h:
            m[r, max(a, 0):min(b, w - 1) + 1] = True
    return m


# ------------------------------------------------------------- region algebra
def test_region_runs_are_canonical(capi, oracle):
    rng = synth.SplitMix64(77)
    for k in range(5):
        raw = synth.random_domain(rng, 40, 70, 60, max_len=12)
        raw[:, 2] = np.maximum(raw[:, 1], raw[:, 2])
        want = synth.mask_runs(mask_of(raw, 40, 70))
        got = capi.Region(raw).runs()
        assert np.array_equal(got, want)
        assert np.array_equal(oracle.domain_runs(raw, 40, 70), want)
    # adjacent runs merge, a gap of one column does not
    assert np.array_equal(capi.Region([(1, 0, 3), (1, 4, 6), (1, 8, 9)]).runs(), [[1, 0, 6], [1, 8, 9]])
    with pytest.raises(capi.MvError) as e:
        capi.Region([(0, 5, 4)])
    assert e.value.status == -1


def test_reduce_domain_clips_and_intersects(capi):
    img = capi.Image(np.zeros((30, 40), np.uint8))
    assert np.array_equal(img.domain().runs(), synth.rect_runs(0, 0, 29, 39))
    a = synth.disk_runs(5.0, 5.0, 8.0)                       # partly outside the image
    b = synth.rect_runs(0, 3, 12, 9)
    ra = img.reduce_domain(a)
    assert np.array_equal(ra.domain().runs(), synth.mask_runs(mask_of(a, 30, 40)))
    rab = ra.reduce_domain(b)
    assert np.array_equal(rab.domain().runs(), synth.mask_runs(mask_of(a, 30, 40) & mask_of(b, 30, 40)))
    assert np.array_equal(rab.full_domain().domain().runs(), synth.rect_runs(0, 0, 29, 39))
    assert len(img.reduce_domain(synth.rect_runs(50, 50, 60, 60)).domain().runs()) == 0
    rect = capi.Region.rectangle1(2.4, 3.6, 5, 7)
    assert np.array_equal(rect.runs(), synth.rect_runs(2, 4, 5, 7))
    out = capi.C.c_void_p()
    with pytest.raises(capi.MvError) as e:                   # a region is not an image
        capi._check(capi.lib().mv_reduce_domain(rect.h, rect.h, capi.C.byref(out)))
    assert e.value.status == -3


# -------------------------------------------------------------- domain create
def test_create_over_rectangle_equals_cropped_template(capi, oracle, scene_small):
    from test_persist import same_rasters
    big = np.zeros((60, 70), np.uint8)
    t = scene_small.template
    big[11:11 + t.shape[0], 17:17 + t.shape[1]] = t
    runs = synth.rect_runs(11, 17, 10 + t.shape[0], 16 + t.shape[1])
    m = capi.NccModel(capi.Image(big, domain=runs), 0, -0.3, 0.6)
    o = oracle.Model(big, levels=0, start=-0.3, extent=0.6, domain=runs)
    ref = oracle.Model(t, levels=0, start=-0.3, extent=0.6)
    same_rasters(m, ref)
    same_oracle_rasters(o, ref)
    assert (o.origin_row, o.origin_col) == (ref.origin_row + 11, ref.origin_col + 17)
    assert (m.origin_row, m.origin_col) == (o.origin_row, o.origin_col)


def ellipse(h, w):
    yy, xx = np.mgrid[0:h, 0:w]
    return ((yy - (h - 1) / 2) / (h / 2)) ** 2 + ((xx - (w - 1) / 2) / (w / 2 - .3)) ** 2 <= 1


def test_create_over_ellipse(capi, oracle, scene_small):
    """Library and oracle agree bit for bit; the oracle's create equals the
    rebuild of a level 0 computed here from the definition."""
    from test_persist import same_rasters
    t = scene_small.template
    h, w = t.shape
    mask = ellipse(h, w)
    rows, cols = np.nonzero(mask)
    assert rows.min() == 0 and rows.max() == h - 1 and cols.min() == 0 and cols.max() == w - 1
    big = np.zeros((h + 9, w + 5), np.uint8)
    big[4:4 + h, 2:2 + w] = t
    runs = synth.mask_runs(mask, 4, 2)
    m = capi.NccModel(capi.Image(big, domain=runs), 0, -0.3, 0.6)
    o = oracle.Model(big, levels=0, start=-0.3, extent=0.6, domain=runs)
    same_rasters(m, o)
    assert m.serialize() == o.serialize()
    # level 0 by hand: gray centred over the mask, zero outside; centroids in long double
    n = int(mask.sum())
    mean = float(int(t[mask].astype(np.int64).sum())) / float(n)
    gray = np.where(mask, t.astype(np.float64) - mean, 0.0)
    ld = np.longdouble
    sr = sum(ld(int(r)) * ld(int(b - a + 1)) for r, a, b in synth.mask_runs(mask))
    sc = sum((ld(int(a)) + ld(int(b))) * ld(int(b - a + 1)) / ld(2) for r, a, b in synth.mask_runs(mask))
    lor, loc = float(sr / ld(n)), float(sc / ld(n))
    sra = sr + ld(4) * ld(n)
    sca = sc + ld(2) * ld(n)
    orow, ocol = float(sra / ld(n)), float(sca / ld(n))
    r17 = lambda v: format(float(v), ".17g")
    body = [f"0 0 {r17(-0.3)} {r17(0.6)} {r17(o.step)}", f"{r17(orow)} {r17(ocol)} {w} {h}", f"{r17(lor)} {r17(loc)}"]
    body += [f"{int(k)} {r17(g)}" for k, g in zip(mask.ravel(), gray.ravel())]
    text = ("\n".join(body) + "\n").encode()
    data = b"MVKIT_NCC 1\n" + str(oracle.checksum(text)).encode() + b"\n" + text
    hand = oracle.Model.deserialize(data)
    same_oracle_rasters(o, hand)
    assert (o.origin_row, o.origin_col) == (hand.origin_row, hand.origin_col)


def test_domain_create_rejections(capi, scene_small):
    t = scene_small.template
    with pytest.raises(capi.MvError) as e:
        capi.NccModel(capi.Image(t, domain=[(0, 0, 2)]), 0, 0.0, 0.0)    # three pixels
    assert e.value.status == -1 and e.value.parameter == "domain"


# ------------------------------------------------------------------ ROI find
@pytest.fixture(scope="module")
def small(oracle, scene_small):
    m = oracle.Model(scene_small.template, levels=0, start=-E, extent=2 * E)
    full = m.find(scene_small.image, -E, 2 * E, 0.7, 0, 0.5)
    assert len(full) == 2
    return m, full


def test_full_rectangle_domain_is_the_full_search(scene_small, small):
    m, full = small
    h, w = scene_small.image.shape
    got = m.find(scene_small.image, -E, 2 * E, 0.7, 0, 0.5, domain=synth.rect_runs(0, 0, h - 1, w - 1))
    assert np.array_equal(got, full)


@pytest.mark.parametrize("shape", ["rect", "disk"])
def test_domain_around_an_instance(scene_small, small, shape):
    m, full = small
    for inst in full:
        cy, cx = inst[0], inst[1]
        if shape == "rect":
            dom = synth.rect_runs(int(cy) - 25, int(cx) - 25, int(cy) + 25, int(cx) + 25)
        else:
            dom = synth.disk_runs(round(cy), round(cx), 3.0)
        got = m.find(scene_small.image, -E, 2 * E, 0.7, 0, 0.5, domain=dom)
        assert got.shape == (1, 4)
        assert np.abs(got[0] - inst).max() < 1e-9


def test_empty_domain_finds_nothing(scene_small, small):
    m, _ = small
    assert len(m.find(scene_small.image, -E, 2 * E, 0.7, 0, 0.5, domain=np.zeros((0, 3)))) == 0


def test_random_sparse_domains(oracle, scene_small, small):
    m, full = small
    h, w = scene_small.image.shape
    rng = synth.SplitMix64(4)
    for k in range(6):
        raw = synth.random_domain(rng, h, w, 40 + 40 * k, max_len=25)
        dom = oracle.domain_runs(raw, h, w)
        mask = mask_of(dom, h, w)
        got = m.find(scene_small.image, -E, 2 * E, 0.3, 0, 0.5, domain=dom)
        for g in got:
            r, c = int(math.floor(g[0] + .5)), int(math.floor(g[1] + .5))
            assert mask[r, c]
            near = [f for f in full if math.hypot(f[0] - g[0], f[1] - g[1]) < 2]
            for f in near:
                assert g[3] <= f[3] + 1e-12


def test_snapped_candidates_stay_in_the_domain(scene_small, small):
    """A band of columns 2-4 px beside an instance's centre: level-1
    candidates snap into it and the fit cannot leave it."""
    m, full = small
    cy, cx = full[0][0], full[0][1]
    c0 = int(round(cx)) + 2
    dom = synth.rect_runs(int(cy) - 15, c0, int(cy) + 15, c0 + 2)
    got = m.find(scene_small.image, -E, 2 * E, 0.3, 0, 0.5, domain=dom)
    assert len(got) >= 1
    for g in got:
        assert c0 <= math.floor(g[1] + .5) <= c0 + 2
    assert got[0][3] < full[0][3]


def test_find_models_full_rectangle_domain(oracle):
    s = synth.small_multi()
    e = math.radians(16)
    models = [oracle.Model(t, levels=3, start=-e, extent=2 * e) for t in s.templates]
    opts = [dict(start=-e, extent=2 * e, min_score=0.7, max_overlap=0.5, num_matches=0)] * 2
    h, w = s.image.shape
    want = oracle.find_models(models, s.image, opts, 0, 0.5)
    got = oracle.find_models(models, s.image, opts, 0, 0.5, domain=synth.rect_runs(0, 0, h - 1, w - 1))
    assert len(want) == 4 and np.array_equal(got, want)
