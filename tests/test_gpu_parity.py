"""GPU parity: the HIP find path against the CPU oracle.

Integer passes are bit-exact; find results must have the oracle's instance
count and order, row/col within 0.02 px, angle within 0.01 degree and score
within 1e-4.  Everything goes through the library's C ABI (mv_* and the
mvamd_* / ncchip_* extensions)."""
import ctypes as C
import json
import math
import os
import sys

import numpy as np
import pytest

from conftest import ANGLE_TOL, POS_TOL, SCORE_TOL

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    torch.cuda.set_device(0)
    return torch.device("cuda:0")


def check_matches(got, want, label=""):
    assert got.shape == want.shape, f"{label}: count {len(got)} != {len(want)}\n{got}\n{want}"
    if len(want) == 0:
        return
    assert np.abs(got[:, 0] - want[:, 0]).max() <= POS_TOL, label
    assert np.abs(got[:, 1] - want[:, 1]).max() <= POS_TOL, label
    assert np.abs(got[:, 2] - want[:, 2]).max() <= ANGLE_TOL, label
    assert np.abs(got[:, 3] - want[:, 3]).max() <= SCORE_TOL, label


# ----------------------------------------------------------- integer passes
def block_sums(img, levels):
    out, cur = [], img.astype(np.int64)
    for _ in range(1, levels):
        h, w = cur.shape
        nh, nw = max(1, h // 2), max(1, w // 2)
        r0 = 2 * np.arange(nh); r1 = np.minimum(r0 + 1, h - 1)
        c0 = 2 * np.arange(nw); c1 = np.minimum(c0 + 1, w - 1)
        cur = cur[r0][:, c0] + cur[r0][:, c1] + cur[r1][:, c0] + cur[r1][:, c1]
        out.append(cur)
    return out


@pytest.mark.parametrize("shape", [(2048, 2048), (480, 640), (37, 53), (1, 9), (3, 1)])
def test_pyramid_bit_exact(capi, oracle, shape):
    say(f"pyramid {shape}")
    rng = np.random.default_rng(11)
    img = rng.integers(0, 256, size=shape, dtype=np.uint8)
    levels = 6
    lib = capi.lib()
    lib.ncchip_pyramid_level.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                         C.c_int, C.c_int, C.c_void_p]
    src = torch.from_numpy(img).cuda()
    kinds = {0: (0, torch.uint8), 1: (1, torch.int16), 2: (2, torch.int32)}
    prev, prev_kind = src, 0
    want = block_sums(img, levels)
    ref = oracle.pyramid(img, levels)
    dc = float(img[0, 0])
    for l in range(1, levels):
        h, w = want[l - 1].shape
        kind = 1 if l <= 4 else 2
        dst = torch.zeros((h, w), dtype=kinds[kind][1], device="cuda")
        st = lib.ncchip_pyramid_level(prev.data_ptr(), prev_kind, prev.shape[1], prev.shape[1], prev.shape[0],
                                      dst.data_ptr(), kind, w, w, h, None)
        assert st == 0
        torch.cuda.synchronize()
        got = dst.cpu().numpy().astype(np.int64)
        if kind == 1:
            got &= 0xFFFF
        assert np.array_equal(got, want[l - 1]), f"level {l}"
        # the reference's FP64 level values are exactly S * 4^-l - dc
        assert np.array_equal(got * 4.0 ** -l - dc, ref[l - 1]), f"level {l} vs oracle"
        prev, prev_kind = dst, kind


def test_row_prefix_and_support_bit_exact(capi, oracle):
    say("row prefix")
    rng = np.random.default_rng(12)
    img = rng.integers(0, 256, size=(64, 300), dtype=np.uint8)
    img[10:20, 30:200] = img[0, 0]  # flat region: zero support
    lib = capi.lib()
    lib.ncchip_row_prefix.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    h, w = img.shape
    src = torch.from_numpy(img).cuda()
    p1 = torch.zeros((h, w + 1), dtype=torch.int64, device="cuda")
    p2 = torch.zeros((h, w + 1), dtype=torch.int64, device="cuda")
    nz = torch.zeros((h, w + 1), dtype=torch.int32, device="cuda")
    sup = torch.zeros((h + 1, w + 1), dtype=torch.int32, device="cuda")
    assert lib.ncchip_row_prefix(src.data_ptr(), 0, w, w, h, src.data_ptr(), 0, p1.data_ptr(), p2.data_ptr(),
                                 nz.data_ptr(), sup.data_ptr(), None) == 0
    torch.cuda.synchronize()
    s = img.astype(np.int64)
    assert np.array_equal(p1.cpu().numpy()[:, 1:].astype(np.int64), np.cumsum(s, axis=1))
    assert np.array_equal(p2.cpu().numpy()[:, 1:], np.cumsum(s * s, axis=1))
    # reference values: prefix of q = S - dc, exact in FP64 (matcher.cpp:487-495)
    ps, pq = oracle.row_prefix(s - float(img[0, 0]))
    P1 = p1.cpu().numpy().astype(np.int64)
    assert np.array_equal(P1 - np.arange(w + 1) * int(img[0, 0]), ps)
    q = (s != img[0, 0]).astype(np.int64)
    integral = np.zeros((h + 1, w + 1), np.int64)
    integral[1:, 1:] = q.cumsum(0).cumsum(1)
    assert np.array_equal(sup.cpu().numpy().astype(np.int64), integral)


# ------------------------------------------------------------------- find
def run_find(capi, oracle, scene, levels, start, extent, min_score, nm, overlap=0.5, subpixel=True, highest=0,
             lowest=1, absolute=False, form=0, label=""):
    say(f"find {label}: {scene.image.shape} levels={levels} extent={extent:.3f} nm={nm} form={form}")
    metric = "ignore_global_polarity" if absolute else "use_polarity"
    m = capi.NccModel(scene.template, levels if levels else "auto", start, extent, "auto", metric)
    o = oracle.Model(scene.template, levels=levels, start=start, extent=extent, absolute=absolute)
    want, tr = o.find(scene.image, start, extent, min_score, nm, overlap, subpixel, highest, lowest, trace=True)
    opts = capi.NccModel.options(start, extent, min_score, nm, overlap, subpixel, highest, lowest, form)
    g = m.find_trace(scene.image, opts)
    return m, want, tr, g


CASES = [
    ("plain640", "scene_plain", 3, 0.0, 0.0, 0.8, 3),
    ("small_rot", "scene_small", 0, -math.radians(20), math.radians(40), 0.7, 2),
    ("small_rot_all", "scene_small", 0, -math.radians(20), math.radians(40), 0.5, 0),
    ("rot2048", "scene_rot", 4, -math.pi / 6, math.pi / 3, 0.7, 5),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_find_matches_oracle(capi, oracle, request, case):
    label, fixture, levels, start, extent, ms, nm = case
    scene = request.getfixturevalue(fixture)
    m, want, tr, g = run_find(capi, oracle, scene, levels, start, extent, ms, nm, label=label)
    check_matches(g["results"], want, label)
    # stage by stage: identical coarse peaks and full-scored candidate list
    assert len(g["coarse"]) == len(tr["coarse"])
    assert np.array_equal(g["coarse"][:, :3], tr["coarse"][:, :3])
    assert np.abs(g["coarse"][:, 3] - tr["coarse"][:, 3]).max() <= 1e-5
    assert len(g["prepared"]) == len(tr["prepared"])
    fin = np.isfinite(tr["prepared"][:, 3])
    assert np.array_equal(np.isfinite(g["prepared"][:, 3]), fin)
    assert np.array_equal(g["prepared"][fin, :3], tr["prepared"][fin, :3])
    assert np.abs(g["prepared"][fin, 3] - tr["prepared"][fin, 3]).max() <= 1e-12
    # every candidate the oracle refined was refined to the same verdict and pose
    ro, rg = tr["refined"], g["refined"]
    done = ro[:, 4] >= 0
    assert np.all(rg[done, 4] >= 0)
    assert np.array_equal(ro[done, 4], rg[done, 4])
    assert np.abs(ro[done, :2] - rg[done, :2]).max() <= POS_TOL
    # the C ABI (HALCON output convention) returns the same matches
    shifted = capi.halcon_shift(want)
    c = m.find(scene.image, start, extent, ms, nm, 0.5, "true", 0)
    check_matches(c, shifted, label + " via C ABI")


def test_truth_recovered_rot2048(capi, scene_rot):
    """All five pasted instances of the 2048x2048 workload are found."""
    m = capi.NccModel(scene_rot.template, 4, -math.pi / 6, math.pi / 3)
    res = m.find_trace(scene_rot.image, capi.NccModel.options(-math.pi / 6, math.pi / 3, 0.7, 5, 0.5))["results"]
    assert len(res) == 5
    for r, c, a, s in res:
        t = min(scene_rot.truth, key=lambda t: math.hypot(r - t[0], c - t[1]))
        assert math.hypot(r - t[0], c - t[1]) < 0.2 and abs(a - t[2]) < 2e-3 and s > 0.9


@pytest.mark.parametrize("variant", ["no_subpixel", "lowest2", "absolute", "highest2", "narrow_query"])
def test_find_variants(capi, oracle, scene_small, variant):
    start, extent = -math.radians(20), math.radians(40)
    kw = dict(levels=0, start=start, extent=extent, min_score=0.6, nm=0)
    qstart, qextent = start, extent
    opts = {}
    if variant == "no_subpixel":
        opts["subpixel"] = False
    elif variant == "lowest2":
        opts["lowest"] = 2
    elif variant == "absolute":
        opts["absolute"] = True
    elif variant == "highest2":
        opts["highest"] = 2
    elif variant == "narrow_query":
        qstart, qextent = -0.1234, 0.3  # endpoints off the trained grid
    metric = "ignore_global_polarity" if opts.get("absolute") else "use_polarity"
    say(f"variant {variant}")
    m = capi.NccModel(scene_small.template, "auto", start, extent, "auto", metric)
    o = oracle.Model(scene_small.template, levels=0, start=start, extent=extent, absolute=opts.get("absolute", False))
    want = o.find(scene_small.image, qstart, qextent, 0.6, 0, 0.5, opts.get("subpixel", True), opts.get("highest", 0),
                  opts.get("lowest", 1))
    got = m.find_trace(scene_small.image, capi.NccModel.options(qstart, qextent, 0.6, 0, 0.5, opts.get("subpixel", True),
                                                                opts.get("highest", 0), opts.get("lowest", 1)))
    check_matches(got["results"], want, variant)


def test_no_candidates_and_out_of_range_query(capi, scene_small):
    say("no candidates")
    m = capi.NccModel(scene_small.template, "auto", 0.0, 0.3)
    flat = np.full((120, 160), 77, np.uint8)
    assert m.find(flat, 0.0, 0.3, 0.5, 0).shape == (0, 4)
    # a query range disjoint from the trained range finds nothing
    assert m.find(scene_small.image, 1.0, 0.2, 0.5, 0).shape == (0, 4)


def test_device_resident_find_matches_host(capi, scene_rot):
    say("device-resident find")
    m = capi.NccModel(scene_rot.template, 4, -math.pi / 6, math.pi / 3)
    opts = capi.NccModel.options(-math.pi / 6, math.pi / 3, 0.7, 5, 0.5)
    img = torch.from_numpy(scene_rot.image).cuda()
    a, _ = m.find_device(img.data_ptr(), 2048, 2048, 2048, opts)
    b = m.find_trace(scene_rot.image, opts)["results"]
    assert np.array_equal(a, b)
    # a pitched view (every other column block) gives the same result as a packed copy
    wide = torch.zeros((2048, 2304), dtype=torch.uint8, device="cuda")
    wide[:, :2048] = img
    c, _ = m.find_device(wide.data_ptr(), 2048, 2048, 2304, opts)
    assert np.array_equal(a, c)


def test_batched_find_equals_single_finds(capi, oracle):
    """A batch of frames of different sizes (and one without matches) returns,
    frame by frame, exactly what single finds return, and matches the oracle."""
    say("batched find: 4 frames")
    from mvkit_amd import synth
    scenes = [synth.small_rot(frame=k) for k in range(3)]
    m = capi.NccModel(scenes[0].template, "auto", -math.radians(20), math.radians(40))
    opts = capi.NccModel.options(-math.radians(20), math.radians(40), 0.6, 0, 0.5)
    imgs = [s.image for s in scenes] + [np.full((100, 140), 90, np.uint8)]
    imgs[1] = np.ascontiguousarray(imgs[1][:170, :230])  # another size
    dev = [torch.from_numpy(i).cuda() for i in imgs]
    h = [i.shape[0] for i in imgs]
    w = [i.shape[1] for i in imgs]
    batch, _ = m.find_batch([d.data_ptr() for d in dev], w, h, w, opts)
    assert len(batch) == len(imgs)
    for k, d in enumerate(dev):
        single, _ = m.find_device(d.data_ptr(), w[k], h[k], w[k], opts)
        assert np.array_equal(batch[k], single), (k, batch[k], single)
    assert len(batch[3]) == 0
    om = oracle.Model(scenes[0].template, levels=0, start=-math.radians(20), extent=math.radians(40))
    for k in range(3):
        want = om.find(imgs[k], -math.radians(20), math.radians(40), 0.6, 0, 0.5)
        check_matches(batch[k], np.asarray(want).reshape(-1, 4), f"frame {k}")


GOLDEN = sorted(p for p in os.listdir(os.path.join(os.path.dirname(__file__), "golden")) if p.endswith(".json"))


@pytest.mark.parametrize("name", GOLDEN)
def test_find_matches_golden(capi, name):
    """The C ABI find against the committed fixtures (oracle results on
    generated scenes whose buffers are pinned by SHA-256)."""
    from mvkit_amd import synth
    say(f"golden {name}")
    with open(os.path.join(os.path.dirname(__file__), "golden", name)) as f:
        g = json.load(f)
    s = getattr(synth, g["scene"]["factory"])(**g["scene"]["kwargs"])
    assert s.sha256() == g["scene"]["sha256"]
    md, fd = g["model"], g["find"]
    m = capi.NccModel(s.template, md["levels"] if md["levels"] else "auto", md["angle_start"], md["angle_extent"])
    got = m.find(s.image, md["angle_start"], md["angle_extent"], fd["min_score"], fd["num_matches"], fd["max_overlap"])
    want = capi.halcon_shift(np.array(g["results"]).reshape(-1, 4))
    check_matches(got, want, name)


FORM_CASES = [
    ("plain640", "scene_plain", 3, 0.0, 0.0, 0.8, 3),
    ("plain640_1level", "scene_plain", 1, 0.0, 0.0, 0.8, 3),
    ("small_rot", "scene_small", 0, -math.radians(20), math.radians(40), 0.7, 2),
    ("small_rot_all", "scene_small", 0, -math.radians(20), math.radians(40), 0.5, 0),
    ("rot2048", "scene_rot", 4, -math.pi / 6, math.pi / 3, 0.7, 5),
]


def coarse_errors(got, want, thr):
    """Score errors of the GPU coarse peaks against the FP64 oracle's at the
    same (row, col, angle); oracle peaks clear of the threshold must be found."""
    index = {(r, c, a): s for r, c, a, s in got}
    errs = []
    for r, c, a, s in want:
        g = index.get((r, c, a))
        if g is None:
            assert s < thr + 2e-4, f"oracle coarse peak {(r, c, a, s)} missing"
            continue
        errs.append(abs(g - s))
    return np.array(errs)


@pytest.mark.parametrize("case", FORM_CASES, ids=[c[0] for c in FORM_CASES])
def test_split_bf16_find_matches_oracle(capi, oracle, request, case):
    """Split-BF16 MFMA coarse correlation: final matches within the parity
    tolerances and every coarse score within 1e-4 of the FP64 oracle."""
    label, fixture, levels, start, extent, ms, nm = case
    scene = request.getfixturevalue(fixture)
    m, want, tr, g = run_find(capi, oracle, scene, levels, start, extent, ms, nm, form=capi.FORM_BF16X3,
                              label=label + " bf16x3")
    check_matches(g["results"], want, label)
    errs = coarse_errors(g["coarse"], tr["coarse"], ms * .4)
    say(f"  bf16x3 coarse score error vs FP64: max {errs.max():.3g} rms {np.sqrt((errs ** 2).mean()):.3g} "
        f"over {len(errs)} peaks")
    assert errs.max() <= 1e-4


def test_coarse_score_error_per_form(capi, oracle, scene_rot):
    """Each correlation form's coarse score error against the FP64 oracle on
    the 2048^2 +-30 degree workload (FP64 form: identical to rounding)."""
    o = oracle.Model(scene_rot.template, levels=4, start=-math.pi / 6, extent=math.pi / 3)
    _, tr = o.find(scene_rot.image, -math.pi / 6, math.pi / 3, 0.7, 5, 0.5, trace=True)
    m = capi.NccModel(scene_rot.template, 4, -math.pi / 6, math.pi / 3)
    report = {}
    for name, form in (("fp64", capi.FORM_FP64), ("fp32", capi.FORM_FP32), ("bf16x3", capi.FORM_BF16X3)):
        opts = capi.NccModel.options(-math.pi / 6, math.pi / 3, 0.7, 5, 0.5, form=form)
        g = m.find_trace(scene_rot.image, opts)
        errs = coarse_errors(g["coarse"], tr["coarse"], 0.28)
        report[name] = (errs.max(), float(np.sqrt((errs ** 2).mean())), len(errs))
        assert errs.max() <= (1e-12 if name == "fp64" else 1e-4), (name, errs.max())
    say("coarse score error vs FP64 oracle (max, rms, peaks): " +
        ", ".join(f"{k}: {v[0]:.3g} / {v[1]:.3g} / {v[2]}" for k, v in report.items()))


def test_fp8_screen_keeps_the_results(capi, scene_small, scene_rot):
    """The FP8-screened form (E4M3 pass, waves under the screen threshold skip
    their split-BF16 products) returns the split-BF16 form's matches on the
    small scene, and on benchmark frames the same instances within the
    parity tolerances; a screen no window fails gives the split-BF16 coarse
    peaks exactly, one every window fails gives none."""
    from mvkit_amd import synth
    say("fp8 screen: small scene")
    m = capi.NccModel(scene_small.template, "auto", -math.radians(20), math.radians(40))
    args = (-math.radians(20), math.radians(40), 0.7, 0, 0.5)
    bf = m.find_trace(scene_small.image, capi.NccModel.options(*args, form=capi.FORM_BF16X3))
    f8 = m.find_trace(scene_small.image, capi.NccModel.options(*args, form=capi.FORM_FP8_SCREEN))
    assert len(bf["results"]) == 2 and np.array_equal(f8["results"], bf["results"])
    every = m.find_trace(scene_small.image, capi.NccModel.options(*args, form=capi.FORM_FP8_SCREEN, screen=-1e300))
    assert np.array_equal(every["coarse"], bf["coarse"]) and np.array_equal(every["results"], bf["results"])
    none = m.find_trace(scene_small.image, capi.NccModel.options(*args, form=capi.FORM_FP8_SCREEN, screen=1e300))
    assert len(none["coarse"]) == 0 and len(none["results"]) == 0
    a, x = -math.pi / 6, math.pi / 3
    mr = capi.NccModel(scene_rot.template, 4, a, x)
    for k in range(3):
        say(f"fp8 screen: 2048x2048 frame {k}")
        img = scene_rot.image if k == 0 else synth.workload_rot_2048(frame=k).image
        want = mr.find_trace(img, capi.NccModel.options(a, x, 0.7, 5, 0.5, form=capi.FORM_BF16X3))
        got = mr.find_trace(img, capi.NccModel.options(a, x, 0.7, 5, 0.5, form=capi.FORM_FP8_SCREEN))
        # the screen drops coarse peaks under it (and can lift a peak whose
        # suppressor was in a skipped wave), so a match may be reached from
        # another candidate: same instances, poses and scores within tolerance
        g, w = got["results"], want["results"]
        assert len(w) == 5 and g.shape == w.shape, (g, w)
        assert np.abs(g[:, :2] - w[:, :2]).max() <= POS_TOL
        assert np.abs(g[:, 2] - w[:, 2]).max() <= ANGLE_TOL
        assert np.abs(g[:, 3] - w[:, 3]).max() <= SCORE_TOL
        assert 0 < len(got["coarse"]) < len(want["coarse"])


def test_fp8_screen_needs_byte_images(capi):
    from mvkit_amd import synth
    s = synth.small_rot16()
    m = capi.NccModel(s.template, "auto", 0.0, 0.3)
    img = torch.from_numpy(np.ascontiguousarray(s.image).view(np.int16)).cuda()
    h, w = s.image.shape
    opts = capi.NccModel.options(0.0, 0.3, 0.5, 0, 0.5, form=capi.FORM_FP8_SCREEN, pixel_type=capi.PIXEL_UINT2)
    with pytest.raises(capi.MvError) as e:
        m.find_device(img.data_ptr(), w, h, 2 * w, opts)
    assert e.value.status == -6


def test_overlapped_batches_match_single_batches(capi):
    """Two submitted batches in flight return what find_batch returns for
    each; a third submission while two are pending is refused."""
    from mvkit_amd import synth
    say("overlapped batches")
    scenes = [synth.small_rot(frame=k) for k in range(4)]
    m = capi.NccModel(scenes[0].template, "auto", -math.radians(20), math.radians(40))
    opts = capi.NccModel.options(-math.radians(20), math.radians(40), 0.6, 0, 0.5)
    dev = [torch.from_numpy(s.image).cuda() for s in scenes]
    h, w = scenes[0].image.shape
    a_ptrs, b_ptrs = [d.data_ptr() for d in dev[:2]], [d.data_ptr() for d in dev[2:]]
    want_a, _ = m.find_batch(a_ptrs, w, h, w, opts)
    want_b, _ = m.find_batch(b_ptrs, w, h, w, opts)
    ha = m.submit_batch(a_ptrs, w, h, w, opts)
    hb = m.submit_batch(b_ptrs, w, h, w, opts)
    with pytest.raises(capi.MvError) as e:
        m.submit_batch(a_ptrs, w, h, w, opts)
    assert e.value.status == -7
    got_a, _ = m.wait_batch(ha)
    hc = m.submit_batch(a_ptrs, w, h, w, opts)  # a slot is free again
    got_b, _ = m.wait_batch(hb)
    got_c, _ = m.wait_batch(hc)
    for got, want in ((got_a, want_a), (got_b, want_b), (got_c, want_a)):
        assert len(got) == len(want)
        for g, x in zip(got, want):
            assert np.array_equal(g, x)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_matrix_tracking_equals_vector_tracking(capi, request, case):
    """Tracking on matrix cores (split-BF16 screen with exact FP64 re-scoring
    of every window that can be the maximum) gives the FP64 vector kernel's
    tracked candidates: identical poses, and full scores of those poses."""
    label, fixture, levels, start, extent, ms, nm = case
    scene = request.getfixturevalue(fixture)
    say(f"matrix vs vector tracking {label}")
    m = capi.NccModel(scene.template, levels if levels else "auto", start, extent)
    fast = m.find_trace(scene.image, capi.NccModel.options(start, extent, ms, nm, 0.5))
    slow = m.find_trace(scene.image, capi.NccModel.options(start, extent, ms, nm, 0.5, vector_tracking=True))
    assert len(fast["prepared"]) == len(slow["prepared"])
    assert np.array_equal(fast["prepared"][:, :3], slow["prepared"][:, :3])
    fin = np.isfinite(slow["prepared"][:, 3])
    assert np.array_equal(np.isfinite(fast["prepared"][:, 3]), fin)
    assert np.array_equal(fast["prepared"][fin, 3], slow["prepared"][fin, 3])
    assert np.array_equal(fast["results"], slow["results"])


@pytest.mark.parametrize("shape", [(187, 251), (129, 255), (191, 131)])
def test_odd_sized_images_match_oracle(capi, oracle, scene_small, shape):
    """Odd image sizes: every pyramid level drops its last odd row and column."""
    h, w = shape
    img = np.ascontiguousarray(scene_small.image[:h, :w])
    start, extent = -math.radians(20), math.radians(40)
    say(f"odd-sized find {shape}")
    m = capi.NccModel(scene_small.template, 0, start, extent)
    o = oracle.Model(scene_small.template, levels=0, start=start, extent=extent)
    want = capi.halcon_shift(o.find(img, start, extent, 0.5, 0, 0.5))
    got = m.find(img, start, extent, 0.5, 0, 0.5, "true", 0)
    check_matches(got, want, f"odd {shape}")
    opts = capi.NccModel.options(start, extent, 0.5, 0, 0.5, form=capi.FORM_BF16X3)
    d = torch.from_numpy(img).cuda()
    got2, _ = m.find_device(d.data_ptr(), w, h, w, opts)
    check_matches(capi.halcon_shift(got2) if len(got2) else got2, want, f"odd bf16 {shape}")
