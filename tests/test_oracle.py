"""Pins the CPU oracle against closed forms, brute force and invariants.

The reference ships no tests or golden vectors for the NCC path and its NCC
sources need a CMake-generated header, so each oracle function is pinned by
mathematics a plausible mistake would violate (see the docstring of each).
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import ANGLE_TOL, POS_TOL


def block_sums(img: np.ndarray, levels: int):
    """Exact integer block sums with the reference's edge clamp."""
    out = []
    cur = img.astype(np.int64)
    for _ in range(1, levels):
        h, w = cur.shape
        nh, nw = max(1, h // 2), max(1, w // 2)
        r0 = 2 * np.arange(nh)
        r1 = np.minimum(r0 + 1, h - 1)
        c0 = 2 * np.arange(nw)
        c1 = np.minimum(c0 + 1, w - 1)
        cur = cur[r0][:, c0] + cur[r0][:, c1] + cur[r1][:, c0] + cur[r1][:, c1]
        out.append(cur)
    return out


@pytest.mark.parametrize("shape", [(480, 640), (37, 53), (1, 9), (9, 1), (2, 2)])
def test_pyramid_closed_form(oracle, shape):
    """Level l equals S_l * 4^-l - dc exactly (matcher.cpp:436-454)."""
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, size=shape, dtype=np.uint8)
    levels = 5
    got = oracle.pyramid(img, levels)
    want = block_sums(img, levels)
    dc = float(img[0, 0])
    for l, (g, s) in enumerate(zip(got, want), start=1):
        assert np.array_equal(g, s * 4.0 ** -l - dc), f"level {l}"


def test_row_prefix_exact(oracle):
    """Row prefixes equal exact cumulative sums (matcher.cpp:487-495)."""
    rng = np.random.default_rng(2)
    img = rng.integers(0, 256, size=(64, 97), dtype=np.uint8)
    lv = oracle.pyramid(img, 3)[1]
    s, q = oracle.row_prefix(lv)
    assert np.array_equal(s[:, 1:], np.cumsum(lv, axis=1))
    assert np.array_equal(q[:, 1:], np.cumsum(lv * lv, axis=1))
    assert np.all(s[:, 0] == 0) and np.all(q[:, 0] == 0)


def test_zncc_textbook(oracle):
    """normalized_correlation equals the Pearson coefficient; constant input -1."""
    rng = np.random.default_rng(3)
    a, b = rng.normal(size=100), rng.normal(size=100)
    assert oracle.zncc(a, b) == pytest.approx(np.corrcoef(a, b)[0, 1], abs=1e-14)
    assert oracle.zncc(a, -a, absolute=True) == pytest.approx(1.0, abs=1e-14)
    assert oracle.zncc(np.ones(10), b[:10]) == -1


def test_auto_step_and_grid(oracle, scene_rot):
    """Auto step max(.01, min(.15, 1/max(1, hypot(w,h)/2))) and
    ceil(extent/step)+1 angles (model.cpp:222-231)."""
    m = oracle.Model(scene_rot.template, levels=4, start=-math.pi / 6, extent=math.pi / 3)
    step = max(.01, min(.15, 1 / max(1., math.hypot(128, 128) / 2)))
    assert m.step == step
    assert m.n_angles == math.ceil((math.pi / 3) / step) + 1
    angles = [m.patch(0, i)["angle"] for i in range(m.n_angles)]
    assert angles[-1] == -math.pi / 6 + math.pi / 3
    assert np.all(np.diff(angles) > 0)


def test_unrotated_raster_is_normalised_template(oracle, scene_plain):
    """The angle-0 raster is the template, centred and scaled to unit energy,
    at grid coordinates with the centroid as origin (model.cpp:66-75, 39-65)."""
    t = scene_plain.template.astype(np.float64)
    m = oracle.Model(scene_plain.template, levels=3)
    p = m.patch(0, 0)
    assert p["orow"] == 31.5 and p["ocol"] == 31.5
    w = t.ravel() - t.mean()
    w /= np.sqrt((w * w).sum())
    assert np.allclose(p["w"], w, atol=1e-15)
    assert np.array_equal(p["rows"], np.repeat(np.arange(64), 64))
    assert abs(p["w"].sum()) < 1e-12 and abs((p["w"] ** 2).sum() - 1) < 1e-12


def test_quarter_turn_is_permutation(oracle):
    """A 90 degree rotation of a square odd-size template is an exact
    permutation of the unrotated raster (bilinear weights are 0/1)."""
    rng = np.random.default_rng(5)
    t = rng.integers(0, 256, size=(15, 15), dtype=np.uint8)
    m = oracle.Model(t, levels=1, start=0, extent=math.pi / 2, step=math.pi / 16)
    p0 = m.patch(0, 0)
    p90 = m.patch(0, m.n_angles - 1)
    assert p90["angle"] == pytest.approx(math.pi / 2)
    # raster (r, c) relative to the centroid samples template (ca*r+sa*c, -sa*r+ca*c) + origin
    w0 = dict(zip(zip(p0["rows"] - 7, p0["cols"] - 7), p0["w"]))
    for r, c, w in zip(p90["rows"], p90["cols"], p90["w"]):
        assert w == pytest.approx(w0[(c, -r)], abs=1e-12)
    assert p90["n"] == p0["n"]


def _window_zncc(level: np.ndarray, patch: dict, r: int, c: int) -> float:
    vals = level[r + patch["rows"], c + patch["cols"]]
    return float(np.corrcoef(vals, patch["w"])[0, 1])


def test_coarse_peaks_brute_force(oracle):
    """Coarse candidates equal a numpy brute force: ZNCC of every anchor,
    then per-angle 3x3 peaks with the reference's tie rule (matcher.cpp:542-594)."""
    from mvkit_amd import synth
    s = synth.scene(96, 80, 24, 24, 1, 0.0, 0.0, 3.0, False, 12345, integer_positions=True)
    m = oracle.Model(s.template, levels=2)
    _, tr = m.find(s.image, 0, 0, 0.5, 0, 0.5, True, trace=True)
    lv = oracle.pyramid(s.image, 2)[0]
    p = m.patch(1, 0)
    H, W = lv.shape
    plane = np.full((H, W), -np.inf)
    for r in range(-p["rmin"], H - p["rmax"]):
        for c in range(-p["cmin"], W - p["cmax"]):
            plane[r, c] = _window_zncc(lv, p, r, c)
    thr = 0.5 * 0.4
    want = []
    for r in range(H):
        for c in range(W):
            v = plane[r, c]
            if not np.isfinite(v) or v < thr:
                continue
            ok = True
            for dr in (-1, 0, 1):
                for dc in (-1, 0, 1):
                    rr, cc = r + dr, c + dc
                    if (dr or dc) and 0 <= rr < H and 0 <= cc < W:
                        o = plane[rr, cc]
                        if o > v + 1e-12 or (abs(o - v) <= 1e-12 and (dr < 0 or (dr == 0 and dc < 0))):
                            ok = False
            if ok:
                want.append((2.0 * (r + p["orow"]) + 0.5, 2.0 * (c + p["ocol"]) + 0.5, v))
    got = tr["coarse"]
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g[0] == w[0] and g[1] == w[1]
        assert g[3] == pytest.approx(w[2], abs=1e-12)


def test_find_recovers_noiseless_pastes(oracle):
    """Noiseless pastes are found at their pasted poses.  Integer, unrotated
    pastes are exact copies of the template, so the pose is exact and the
    score is 1; rotated pastes are resampled, so they are found within a
    tenth of a pixel.  This pins the whole chain: pyramid, coarse search,
    tracking, continuous score and Gauss-Newton fit."""
    from mvkit_amd import synth
    s = synth.scene(320, 256, 48, 48, 2, 0, 0, 0.0, False, 777, integer_positions=True)
    m = oracle.Model(s.template, levels=3)
    res = m.find(s.image, 0, 0, 0.7, 2, 0.5)
    got = np.array(sorted(map(tuple, res[:, :3])))
    assert np.allclose(got, np.array(sorted(s.truth)), atol=1e-9, rtol=0)
    assert np.allclose(res[:, 3], 1.0, atol=1e-12)
    s = synth.scene(320, 256, 48, 48, 2, -0.3, 0.3, 0.0, False, 777, min_separation=80)
    m = oracle.Model(s.template, levels=3, start=-0.35, extent=0.7)
    res = m.find(s.image, -0.35, 0.7, 0.7, 2, 0.5)
    assert len(res) == 2
    for r, c, a, sc in res:
        t = min(s.truth, key=lambda t: math.hypot(r - t[0], c - t[1]))
        assert math.hypot(r - t[0], c - t[1]) < 0.1 and abs(a - t[2]) < 2e-3 and sc > 0.9


def test_refined_pose_is_local_maximum(oracle, scene_small):
    """First-order optimality of the GN fit: small steps in row, column or
    angle do not raise the continuous score beyond the fit's stop tolerance."""
    s = scene_small
    m = oracle.Model(s.template, levels=0, start=-math.radians(20), extent=math.radians(40))
    res = m.find(s.image, -math.radians(20), math.radians(40), 0.7, 2, 0.5)
    assert len(res) >= 1
    for r, c, a, sc in res:
        base = m.continuous(s.image, r, c, a)
        assert base == pytest.approx(sc, abs=1e-12)
        for dr, dc, da in [(0.02, 0, 0), (-0.02, 0, 0), (0, 0.02, 0), (0, -0.02, 0), (0, 0, 2e-3), (0, 0, -2e-3)]:
            assert m.continuous(s.image, r + dr, c + dc, a + da) <= base + 2e-5


def test_overlap_closed_forms(oracle):
    """Overlap = intersection area / template area (matcher.cpp:252-277)."""
    w, h, orow, ocol = 40, 20, 9.5, 19.5
    assert oracle.overlap((50, 50, 0), (50, 50, 0), w, h, orow, ocol) == pytest.approx(1)
    assert oracle.overlap((50, 50, 0), (50, 60, 0), w, h, orow, ocol) == pytest.approx(30 / 40)
    assert oracle.overlap((50, 50, 0), (55, 60, 0), w, h, orow, ocol) == pytest.approx(30 * 15 / 800)
    assert oracle.overlap((50, 50, 0), (50, 150, 0), w, h, orow, ocol) == 0
    # rotated: Monte Carlo of the rasterised intersection
    rng = np.random.default_rng(7)
    a, b = (50.0, 50.0, 0.3), (54.0, 58.0, -0.4)
    pts = rng.uniform([20, 10], [80, 100], size=(200000, 2))

    def inside(p, pose):
        r, c, t = pose
        dy, dx = p[:, 0] - r, p[:, 1] - c
        u = math.cos(t) * dx - math.sin(t) * dy  # template column offset
        v = math.sin(t) * dx + math.cos(t) * dy  # template row offset
        return (u >= -ocol - .5) & (u <= w - ocol - .5) & (v >= -orow - .5) & (v <= h - orow - .5)

    frac = (inside(pts, a) & inside(pts, b)).mean() * 60 * 90 / (w * h)
    assert oracle.overlap(a, b, w, h, orow, ocol) == pytest.approx(frac, abs=0.02)


def test_failures(oracle):
    """The reference's failure statuses (model.cpp:193-194, 237; matcher.cpp:591)."""
    with pytest.raises(oracle.OracleError) as e:
        oracle.Model(np.full((32, 32), 7, np.uint8))
    assert e.value.status == -1
    from mvkit_amd import synth
    t = synth.make_template(synth.SplitMix64(1), 256, 256)
    with pytest.raises(oracle.OracleError) as e:
        oracle.Model(t, levels=5, start=0, extent=2 * math.pi)
    assert e.value.status == -103
    rng = np.random.default_rng(9)
    img = rng.integers(0, 256, size=(400, 400), dtype=np.uint8)
    m = oracle.Model(img[:16, :16], levels=1)
    with pytest.raises(oracle.OracleError) as e:
        m.find(img, 0, 0, 0.0, 0, 1.0)
    assert e.value.status == -103


# ------------------------------------------------------------ golden fixtures
GOLDEN = sorted(p for p in os.listdir(os.path.join(os.path.dirname(__file__), "golden")) if p.endswith(".json"))


def load_golden(name):
    with open(os.path.join(os.path.dirname(__file__), "golden", name)) as f:
        return json.load(f)


def golden_scene(g):
    from mvkit_amd import synth
    s = getattr(synth, g["scene"]["factory"])(**g["scene"]["kwargs"])
    assert s.sha256() == g["scene"]["sha256"], "synthetic generator drifted"
    return s


@pytest.mark.parametrize("name", [n for n in GOLDEN if "2048" not in n])
def test_oracle_reproduces_golden(oracle, name):
    g = load_golden(name)
    s = golden_scene(g)
    m = oracle.Model(s.template, levels=g["model"]["levels"], start=g["model"]["angle_start"],
                     extent=g["model"]["angle_extent"])
    f = g["find"]
    res, tr = m.find(s.image, g["model"]["angle_start"], g["model"]["angle_extent"], f["min_score"],
                     f["num_matches"], f["max_overlap"], trace=True)
    assert tr["info"].n_coarse == g["n_coarse"] and tr["info"].n_prepared == g["n_prepared"]
    assert np.array_equal(res, np.array(g["results"]).reshape(-1, 4))
