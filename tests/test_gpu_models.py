"""GPU parity of find_ncc_models: the HIP path (through mv_find_ncc_models and
mvamd_find_ncc_models_device) against the oracle's find_models, for every
form of the per-model controls, repeated handles, per-model level pairs and
the background guard; and finds with models loaded from MVKIT_NCC 1 files,
including a template with a masked (elliptical) domain."""
import math
import sys

import numpy as np
import pytest

from conftest import ANGLE_TOL, POS_TOL, SCORE_TOL

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

E = math.radians(16)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    torch.cuda.set_device(0)
    return torch.device("cuda:0")


@pytest.fixture(scope="module")
def multi():
    from mvkit_amd import synth
    return synth.small_multi()


def check(got, want, label):
    assert got.shape == want.shape, f"{label}: {len(got)} != {len(want)}\n{got}\n{want}"
    if not len(want):
        return
    assert np.array_equal(got[:, 4], want[:, 4]), f"{label}: model order\n{got}\n{want}"
    assert np.abs(got[:, 0] - want[:, 0]).max() <= POS_TOL, label
    assert np.abs(got[:, 1] - want[:, 1]).max() <= POS_TOL, label
    assert np.abs(got[:, 2] - want[:, 2]).max() <= ANGLE_TOL, label
    assert np.abs(got[:, 3] - want[:, 3]).max() <= SCORE_TOL, label


def per_model(v, n):
    return list(v) if isinstance(v, (list, tuple)) else [v] * n


def oracle_opts(n, min_score, num_matches, max_overlap, levels, start=-E, extent=2 * E):
    ms, nm, ov = per_model(min_score, n), per_model(num_matches, n), per_model(max_overlap, n)
    # NumLevels forms in the C ABI's order (ncc_matching_api.cpp:150-153)
    lv = list(levels) if isinstance(levels, (list, tuple)) else [levels]
    if len(lv) == n and n > 1:
        hl = [(h, 1) for h in lv]
    elif len(lv) == 2 * n:
        hl = [(lv[2 * i], lv[2 * i + 1]) for i in range(n)]
    elif len(lv) == 2:
        hl = [tuple(lv)] * n
    else:
        hl = [(lv[0], 1)] * n
    return [dict(start=start, extent=extent, min_score=ms[i], max_overlap=ov[i], num_matches=nm[i], subpixel=True,
                 highest=hl[i][0], lowest=hl[i][1]) for i in range(n)]


def selection(n, num_matches, max_overlap):
    sn = num_matches if not isinstance(num_matches, (list, tuple)) else -1
    so = max_overlap if not isinstance(max_overlap, (list, tuple)) else -1.0
    return sn, so


def run(capi, oracle, image, cmodels, omodels, min_score, num_matches, max_overlap, levels=0, label="", start=-E,
        extent=2 * E):
    n = len(cmodels)
    say(f"find_models {label}: {image.shape} models={n} nm={num_matches} overlap={max_overlap} levels={levels}")
    got = capi.find_models(cmodels, image, start, extent, min_score, num_matches, max_overlap, "true", levels)
    sn, so = selection(n, num_matches, max_overlap)
    want = oracle.find_models(omodels, image,
                              oracle_opts(n, min_score, num_matches, max_overlap, levels, start, extent), sn, so)
    if len(want):
        want[:, :4] = capi.halcon_shift(want[:, :4])
    check(got, want, label)
    return got, want


@pytest.fixture(scope="module")
def models(capi, oracle, multi):
    c = [capi.NccModel(t, 3, -E, 2 * E) for t in multi.templates]
    o = [oracle.Model(t, levels=3, start=-E, extent=2 * E) for t in multi.templates]
    twin_c = capi.NccModel(multi.templates[0], 3, -E, 2 * E)
    twin_o = oracle.Model(multi.templates[0], levels=3, start=-E, extent=2 * E)
    return c, o, twin_c, twin_o


CASES = [
    ("global", 0.7, 0, 0.5, 0),
    ("global_quota1", 0.7, 1, 0.5, 0),
    ("per_model", 0.7, [2, 2], [0.5, 0.5], 0),
    ("per_model_scores", [0.6, 0.8], [0, 1], [0.5, 0.3], [3, 3]),
    ("level_pairs", 0.7, 0, 0.5, [0, 1, 0, 2]),
    ("per_model_highest", 0.7, 0, 0.5, [3, 2]),
    ("loose", 0.5, 0, 0.9, 0),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_find_models_matches_oracle(capi, oracle, multi, models, case):
    label, ms, nm, ov, lv = case
    c, o, _, _ = models
    got, _ = run(capi, oracle, multi.image, c, o, ms, nm, ov, lv, label)
    if label == "per_model":
        assert len(got) == len(multi.truth)


def test_twin_models_cross_overlap_and_two_stage(capi, oracle, multi, models):
    c, o, twin_c, twin_o = models
    run(capi, oracle, multi.image, [c[0], twin_c], [o[0], twin_o], 0.7, 0, 0.5, 0, "twin_global")
    run(capi, oracle, multi.image, [c[0], twin_c], [o[0], twin_o], 0.7, 0, [0.5, 0.5], 0, "twin_per_model")
    got, _ = run(capi, oracle, multi.image, [c[0], twin_c], [o[0], twin_o], 0.7, [1, 1], 0.5, 0, "twin_two_stage")
    assert len(got) == 1


def test_repeated_handle(capi, oracle, multi, models):
    c, o, _, _ = models
    got, _ = run(capi, oracle, multi.image, [c[1], c[1]], [o[1], o[1]], 0.7, [0, 0], [0.5, 0.5], 0, "repeated")
    assert np.array_equal(got[0::2, :4], got[1::2, :4])


def test_background_guard(capi, oracle, multi):
    """Distinct models, one-value NumMatches 0, MinScore in [0.75, 0.85],
    three levels, an angle step <= 0.02 and a query extent <= 0.4 rad: the
    default background guard of the fit is on."""
    a, x = -0.19, 0.38
    c = [capi.NccModel(t, 3, a, x, 0.015) for t in multi.templates]
    o = [oracle.Model(t, levels=3, start=a, extent=x, step=0.015) for t in multi.templates]
    for ms in (0.8, 0.76):
        run(capi, oracle, multi.image, c, o, ms, 0, 0.5, 0, f"guard_{ms}", a, x)


def test_device_entry_matches_host_entry(capi, multi, models):
    c, _, _, _ = models
    img = torch.from_numpy(multi.image).cuda()
    h, w = multi.image.shape
    opts = [capi.NccModel.options(-E, 2 * E, 0.7, 2, 0.5) for _ in c]
    dev = capi.find_models_device(c, img.data_ptr(), w, h, w, opts, -1, -1.0)
    host = capi.find_models(c, multi.image, -E, 2 * E, 0.7, [2, 2], [0.5, 0.5], "true", 0)
    assert dev.shape == host.shape
    assert np.array_equal(dev[:, 4], host[:, 4])
    assert np.allclose(capi.halcon_shift(dev[:, :4]), host[:, :4], atol=1e-12, rtol=0)


def test_find_models_rot2048(capi, oracle, scene_rot):
    """The 2048x2048 workload with its template and a second, absent one."""
    from mvkit_amd import synth
    other = synth.make_template(synth.SplitMix64(12345), 96, 80)
    a, x = -math.pi / 6, math.pi / 3
    c = [capi.NccModel(scene_rot.template, 4, a, x), capi.NccModel(other, 4, a, x)]
    o = [oracle.Model(scene_rot.template, levels=4, start=a, extent=x), oracle.Model(other, levels=4, start=a, extent=x)]
    say("find_models rot2048")
    got = capi.find_models(c, scene_rot.image, a, x, 0.7, 0, 0.5, "true", 0)
    want = oracle.find_models(o, scene_rot.image, [dict(start=a, extent=x, min_score=0.7, max_overlap=0.5,
                                                        num_matches=0)] * 2, 0, 0.5)
    want[:, :4] = capi.halcon_shift(want[:, :4])
    check(got, want, "rot2048")
    assert np.count_nonzero(got[:, 4] == 0) == 5


def test_loaded_models_find_like_created(capi, oracle, scene_small, tmp_path):
    a, x = -math.radians(20), math.radians(40)
    m = capi.NccModel(scene_small.template, 0, a, x)
    path = str(tmp_path / "small.ncm")
    m.write(path)
    r = capi.NccModel.read(path)
    say("find with a model read from a file")
    want = m.find(scene_small.image, a, x, 0.7, 2, 0.5, "true", 0)
    got = r.find(scene_small.image, a, x, 0.7, 2, 0.5, "true", 0)
    assert len(want) == 2 and np.array_equal(got, want)


def test_masked_template_model_matches_oracle(capi, oracle):
    """A reduced-domain (elliptical) template as the reference writes it:
    loaded into both, its pastes are found with the oracle's poses."""
    from mvkit_amd import synth
    from test_persist import masked_file
    data, mask = masked_file(oracle)
    t = synth.make_template(synth.SplitMix64(99), 26, 30)
    rng = synth.SplitMix64(4242)
    base = synth.blurred_noise(rng, 200, 240).astype(np.float64)
    for cy, cx, ang in [(60.3, 70.8, 0.12), (140.6, 170.2, -0.2)]:
        synth.paste(base, t, cy, cx, ang)
    image = np.rint(np.clip(base + 3.0 * rng.normal(200 * 240).reshape(200, 240), 0, 255)).astype(np.uint8)
    o = oracle.Model.deserialize(data)
    m = capi.NccModel.deserialize(data)
    say("find with a masked-template model")
    want = o.find(image, -0.25, 0.5, 0.6, 2, 0.5)
    assert len(want) == 2
    got = m.find(image, -0.25, 0.5, 0.6, 2, 0.5, "true", 0)
    shifted = capi.halcon_shift(want)
    assert got.shape == shifted.shape
    assert np.abs(got[:, :2] - shifted[:, :2]).max() <= POS_TOL
    assert np.abs(got[:, 2] - shifted[:, 2]).max() <= ANGLE_TOL
    assert np.abs(got[:, 3] - shifted[:, 3]).max() <= SCORE_TOL
